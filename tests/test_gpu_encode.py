"""GPU encoder vs the CPU oracle, bit for bit, over every mode combination.

Each case prints its name before it touches the GPU so that a failing run
names it.
"""
import itertools
import sys

import numpy as np
import pytest

import imgutil

pytestmark = pytest.mark.gpu

SIZES = [(16, 16), (13, 9), (5, 7), (1, 1), (2, 3), (4, 4), (7, 1), (37, 23)]
NRANDOM = [-1, 0, 5, 64]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, seed, tag, src_kind=0):
    _say(f"case {tag} {w}x{h} dxt={dxt} cd={cd} nrandom={nr} refine={refine} seed={seed}")
    got = gpu.encode_image(q, w, h, dxt, cd, nr, refine, seed, src_kind=src_kind)
    exp = oracle.encode_quantized(q if src_kind == 0 else None, w, h, dxt, cd, nr, refine, seed) \
        if src_kind == 0 else None
    if exp is None:
        raise AssertionError("use _compare_raw for raw sources")
    if not np.array_equal(got, exp):
        bad = np.nonzero(got != exp)[0]
        bs = 8 if dxt == 0 else 16
        blk = bad[0] // bs
        raise AssertionError(f"{tag}: {len(bad)} bytes differ; first block {blk}: "
                             f"gpu {got[blk*bs:(blk+1)*bs].tolist()} oracle {exp[blk*bs:(blk+1)*bs].tolist()}")


@pytest.mark.parametrize("cd", range(8))
def test_all_modes_random(gpu, oracle, cd):
    rng = np.random.default_rng(100 + cd)
    for (w, h), dxt in itertools.product(SIZES[:4], range(3)):
        for kind in ("random", "smooth", "speckle", "extreme"):
            q = imgutil.random_quantized(rng, w, h, dxt, kind)
            for nr, refine in itertools.product(NRANDOM, range(3)):
                _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, 7 + cd, f"{kind}")


@pytest.mark.parametrize("cd", range(8))
def test_edge_sizes(gpu, oracle, cd):
    rng = np.random.default_rng(200 + cd)
    for (w, h), dxt in itertools.product(SIZES[3:], range(3)):
        q = imgutil.random_quantized(rng, w, h, dxt, "smooth")
        for nr, refine in itertools.product([-1, 0, 3], range(3)):
            _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, 3, "edge")


@pytest.mark.parametrize("dxt", range(3))
def test_special_blocks(gpu, oracle, dxt):
    q, w, h = imgutil.special_blocks(dxt)
    for cd, nr, refine in itertools.product(range(8), [-1, 0, 1, 16], range(3)):
        _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, 11, "special")


@pytest.mark.parametrize("cd", range(8))
def test_one_colour_blocks(gpu, oracle, cd):
    """Blocks with one colour (and alpha) among their candidates skip the pair
    scans: whole waves and workgroups of them (top half), mixed with other
    blocks (bottom half), partial blocks at the right and bottom edges."""
    rng = np.random.default_rng(300 + cd)
    for (w, h), dxt in itertools.product([(256, 64), (37, 23)], range(3)):
        q = imgutil.flat_blocks(rng, w, h, dxt)
        for nr, refine in itertools.product([-1, 0, 5, 64], range(3)):
            _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, 5 + cd, "flat")


@pytest.mark.parametrize("cd", range(8))
def test_random_candidates_small_boxes(gpu, oracle, cd):
    """Random candidates drawn from small colour boxes repeat colours: the
    scan over the list without repeats (first occurrences and each repeated
    colour's first pair) must pick the reference's pair, at candidate counts
    where the compacted list is short (m = 16 + 5 .. 16 + 200), with ties."""
    rng = np.random.default_rng(400 + cd)
    for (w, h), dxt in itertools.product([(64, 32), (21, 10)], range(3)):
        q = imgutil.small_box_blocks(rng, w, h, dxt)
        for nr, refine in itertools.product([5, 64, 200], [0, 2]):
            _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, 9 + cd, "smallbox")


def _opaque(q, dxt):
    q = q.reshape(-1, 4).copy()
    q[:, 3] = (1 << imgutil.ALPHABITS[dxt]) - 1
    return q.reshape(-1)


@pytest.mark.parametrize("cd", [0, 1, 3])
def test_random_candidates_saturated_keys(gpu, oracle, cd):
    """The pair scan of the unbounded distances (RGB, YUV, SRGB_MIXED) keys
    its sums saturated at a cap (2^28 - 2; no block of 565 colours with 64
    random candidates has come near it), and a block whose least pair sum
    reaches the cap scans again with exact 64-bit keys.  With the cap set to
    the median least sum of the image (computed here on the CPU), half of the
    blocks take each path, through the plain scan (wide colour boxes) and the
    scan over the list without repeated colours (small boxes)."""
    import ctypes
    L = gpu.lib()
    L.s2tc_amd_debug_set_scan_cap.argtypes = [ctypes.c_uint]
    rng = np.random.default_rng(500 + cd)
    w, h, nr, seed = 32, 16, 64, 13
    try:
        for kind, dxt in itertools.product(["random", "smallbox", "corners"], range(3)):
            if kind == "smallbox":
                q = _opaque(imgutil.small_box_blocks(rng, w, h, dxt), dxt)
            elif kind == "corners":
                q = imgutil.corner_blocks(rng, w, h, dxt)
            else:
                q = _opaque(imgutil.random_quantized(rng, w, h, dxt, "random"), dxt)
            sums = sorted(imgutil.least_pair_sums(oracle, q, w, h, dxt, cd, nr, seed))
            cap = sums[len(sums) // 2]
            assert sums[0] < cap <= sums[-1], f"{kind}: least sums do not straddle the cap"
            _say(f"scan cap {cap} ({kind}, dxt={dxt}, cd={cd})")
            gpu.check(L.s2tc_amd_debug_set_scan_cap(cap))
            for refine in range(3):
                _compare(gpu, oracle, q, w, h, dxt, cd, nr, refine, seed, f"cap-{kind}")
    finally:
        gpu.check(L.s2tc_amd_debug_set_scan_cap(0xffffffff))


def test_raw_rgba8_source(gpu, oracle):
    """Fused DITHER_NONE quantization on load equals rgb565_image + encode."""
    rng = np.random.default_rng(5)
    for (w, h), dxt in itertools.product([(64, 32), (37, 23), (3, 5)], range(3)):
        rgba = imgutil.smooth_rgba8(rng, w, h)
        q = oracle.rgb565_image(rgba, w, h, 4, imgutil.ALPHABITS[dxt], 0)
        for cd, nr, refine in [(5, -1, 1), (3, 0, 2), (7, 8, 2), (2, -1, 0), (0, -1, 2)]:
            _say(f"case raw {w}x{h} dxt={dxt} cd={cd} nrandom={nr} refine={refine}")
            got = gpu.encode_image(rgba, w, h, dxt, cd, nr, refine, 1, src_kind=gpu.SRC_RGBA8)
            exp = oracle.encode_quantized(q, w, h, dxt, cd, nr, refine, 1)
            assert np.array_equal(got, exp), f"raw source mismatch {w}x{h} dxt={dxt} cd={cd}"


def _rgba8_kinds(rng, w, h):
    """RGBA8 images that stress the quick-mode paths: smooth content, uniform
    noise (alpha on both sides of 128, so DXT1 blocks mix transparent and
    opaque texels), flat blocks (all texels equal: equal endpoints), and
    black/white/255-alpha extremes."""
    yield "smooth", imgutil.smooth_rgba8(rng, w, h)
    yield "noise", rng.integers(0, 256, w * h * 4).astype(np.uint8)
    flat = np.empty((h, w, 4), np.uint8)
    flat[...] = rng.integers(0, 256, 4).astype(np.uint8)
    yield "flat", flat.reshape(-1)
    ext = rng.choice(np.array([0, 255, 127, 128], np.uint8), w * h * 4)
    yield "extreme", ext


@pytest.mark.parametrize("cd", [4, 5, 6])
def test_quick_quadratic_modes(gpu, oracle, cd):
    """AVG, W0AVG and WAVG quick mode (dot-product form) on RGBA8 and on
    quantized sources, every refine mode and format, full and partial blocks."""
    rng = np.random.default_rng(300 + cd)
    for (w, h), dxt in itertools.product([(32, 16), (13, 9), (6, 3), (1, 1)], range(3)):
        for kind, rgba in _rgba8_kinds(rng, w, h):
            q = oracle.rgb565_image(rgba, w, h, 4, imgutil.ALPHABITS[dxt], 0)
            for refine in range(3):
                _say(f"case quad {kind} {w}x{h} dxt={dxt} cd={cd} refine={refine}")
                exp = oracle.encode_quantized(q, w, h, dxt, cd, -1, refine, 0)
                got = gpu.encode_image(rgba, w, h, dxt, cd, -1, refine, 0, src_kind=gpu.SRC_RGBA8)
                assert np.array_equal(got, exp), f"raw {kind} {w}x{h} dxt={dxt} cd={cd} refine={refine}"
                got = gpu.encode_image(q, w, h, dxt, cd, -1, refine, 0)
                assert np.array_equal(got, exp), f"quantized {kind} {w}x{h} dxt={dxt} cd={cd} refine={refine}"


@pytest.mark.parametrize("raw", [False, True])
def test_uniform_alpha_blocks(gpu, oracle, raw):
    """DXT5 blocks whose alphas are all equal take a closed form for the
    alpha half: every alpha 0..255 in full blocks, next to blocks that differ
    in one texel's alpha and to partial blocks in the same waves, in quick
    mode (quadratic and generic distances) and with all 16 candidates."""
    rng = np.random.default_rng(55)
    w, h = 4 * 256 + 3, 4 * 3 + 2
    rgba = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    alpha = np.repeat(np.arange(256, dtype=np.uint8), 4)
    for r in range(h):
        rgba[r, :1024, 3] = alpha
    rgba[4:8, 5::9, 3] ^= 1          # one texel off in some blocks of the second block row
    rgba[8:12, :1024:4, 3] = 255     # the third block row: many opaque blocks, a few not
    rgba[8:12, 40::37, 3] = 17
    for refine in range(3):
        for cd, nrandom in [(5, -1), (4, -1), (2, -1), (7, -1), (3, 0), (7, 0)]:
            _say(f"case uniform alpha raw={raw} refine={refine} cd={cd} nrandom={nrandom}")
            if raw:
                got = gpu.encode_image(rgba.reshape(-1), w, h, 2, cd, nrandom, refine, src_kind=gpu.SRC_RGBA8)
                q = oracle.rgb565_image(rgba.reshape(-1), w, h, 4, 8, 0)
            else:
                q = oracle.rgb565_image(rgba.reshape(-1), w, h, 4, 8, 0)
                got = gpu.encode_image(q, w, h, 2, cd, nrandom, refine)
            exp = oracle.encode_quantized(q, w, h, 2, cd, nrandom, refine, 0)
            assert np.array_equal(got, exp), (raw, refine, cd, nrandom)


@pytest.mark.parametrize("cd", [4, 5, 6, 7, 2])
def test_random_candidate_counts(gpu, oracle, cd):
    """Random-candidate blocks with candidate counts 16 + nrandom that are not
    multiples of 16 (odd nrandom; DXT1 blocks whose transparent texels leave
    fewer than 16 colours), up to the LDS-resident limit of 208, every format: the
    bounded distances (AVG, WAVG, W0AVG, NORMALMAP) and DXT5 alpha take the
    keyed pair scan, SRGB the per-step scan with its negative-sum rule."""
    rng = np.random.default_rng(300 + cd)
    for dxt in range(3):
        for nr in (1, 15, 17, 31, 48, 144, 208):
            w, h = (9, 6) if nr < 100 else (5, 4)
            for kind in ("speckle", "random", "extreme"):
                q = imgutil.random_quantized(rng, w, h, dxt, kind)
                _compare(gpu, oracle, q, w, h, dxt, cd, nr, 2 if nr % 2 else 1, 11 + nr, f"candidates-{kind}")


@pytest.mark.parametrize("cd", [7, 2, 0, 5])
def test_wide_candidate_counts(gpu, oracle, cd):
    """More random candidates than LDS holds (nrandom 209 .. 700): the same
    encoder with its candidates, distance rows and column features in device
    scratch, on images of several workgroups' worth of blocks (the persistent
    grid strides over them), every format; NORMALMAP and WAVG take the keyed
    scan, RGB its 64-bit keys, SRGB the per-step scan with its negative-sum
    rule."""
    rng = np.random.default_rng(900 + cd)
    for dxt in range(3):
        for nr, (w, h) in ((209, (23, 17)), (255, (9, 6)), (497, (7, 5)), (700, (5, 4))):
            q = imgutil.random_quantized(rng, w, h, dxt, "random" if nr % 2 else "speckle")
            _compare(gpu, oracle, q, w, h, dxt, cd, nr, 2 if nr % 2 else 1, 3 + nr, "wide")


def test_wide_grid_strides(gpu, oracle):
    """A 264x260 image at nrandom 209 has 1,089 four-block tiles, more than
    the wide kernel's grid of 4 workgroups per CU: workgroups take several
    tiles each and reuse their scratch slice."""
    rng = np.random.default_rng(77)
    w, h = 264, 260
    q = imgutil.random_quantized(rng, w, h, 0, "random")
    _compare(gpu, oracle, q, w, h, 0, 7, 209, 2, 9, "wide-strided")
