"""The drop-in boundary on the GPU against the compiled reference's outputs
(tests/golden/*.npz, written by tests/golden/make_golden.py).

Covers tx_compress_dxtn (environment configuration, dithers, formats, row
strides), rgb565_image (all dithers / srccomps / alphabits), the per-block
s2tc_encode_block_func entry points, and concurrent callers.
"""
import ctypes
import os
import sys
import threading

import numpy as np
import pytest

import imgutil

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["RGB", "YUV", "SRGB", "SRGB_MIXED", "AVG", "WAVG", "W0AVG", "NORMALMAP"]
REFINES = ["NEVER", "ALWAYS", "LOOP"]
DITHERS = ["NONE", "SIMPLE", "FLOYDSTEINBERG"]
ENV_KEYS = ("S2TC_DITHER_MODE", "S2TC_COLORDIST_MODE", "S2TC_RANDOM_COLORS", "S2TC_REFINE_COLORS",
            "S2TC_RANDOM_SEED")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture
def clean_env():
    saved = {k: os.environ.get(k) for k in ENV_KEYS}
    yield
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def test_encoder_matches_golden_blocks(gpu):
    z = load("blocks.npz")
    n = 0
    for key in z.files:
        if not key.startswith("out_"):
            continue
        _, dxt, i, cd, nr, refine = key.split("_")
        dxt, i, cd, nr, refine = int(dxt), int(i), int(cd), int(nr), int(refine)
        w, h = (int(v) for v in z[f"wh_{dxt}_{i}"])
        if n % 100 == 0:
            _say(f"golden block case {key} ({w}x{h})")
        got = gpu.encode_image(z[f"in_{dxt}_{i}"], w, h, dxt, cd, nr, refine, seed=1000 + i)
        assert np.array_equal(got, z[key]), key
        n += 1
    assert n == 2304


def test_rgb565_image_matches_golden(gpu):
    L = gpu.lib()
    L.rgb565_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int]
    z = load("dither.npz")
    for key in z.files:
        if not key.startswith("q_"):
            continue
        _, w, h, sc, ab, d = key.split("_")
        w, h, sc, ab, d = map(int, (w, h, sc, ab, d))
        _say(f"rgb565_image {key}")
        src = z[f"src_{w}_{h}_{sc}"]
        out = np.zeros(w * h * 4, np.uint8)
        L.rgb565_image(out.ctypes.data, src.ctypes.data, w, h, sc, ab, d)
        assert np.array_equal(out, z[key]), key


def test_tx_compress_dxtn_matches_golden(gpu, clean_env):
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    z = load("compress.npz")
    for n, key in enumerate(z["cases"]):
        w, h, fmt, stride, d, cd, nr, refine = map(int, str(key).split("_"))
        os.environ.update(S2TC_DITHER_MODE=DITHERS[d], S2TC_COLORDIST_MODE=NAMES[cd], S2TC_RANDOM_COLORS=str(nr),
                          S2TC_REFINE_COLORS=REFINES[refine], S2TC_RANDOM_SEED="77")
        if n % 50 == 0:
            _say(f"tx_compress_dxtn {key}")
        exp = z["dst_" + str(key)]
        dst = np.full(exp.size, 0xAB, np.uint8)
        src = z[f"src_{w}_{h}"]
        L.tx_compress_dxtn(4, w, h, src.ctypes.data, fmt, dst.ctypes.data, stride)
        assert np.array_equal(dst, exp), key


def test_dither_large_against_oracle(gpu, oracle):
    """SIMPLE and FS on images large enough to span many scan chunks / row bands."""
    L = gpu.lib()
    L.rgb565_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5
    rng = np.random.default_rng(8)
    for (w, h) in [(1000, 700), (257, 513), (4096, 3)]:
        for sc in (3, 4):
            src = imgutil.smooth_rgba8(rng, w, h).reshape(-1, 4)[:, :sc].copy().reshape(-1)
            for ab in (1, 4, 8):
                for d in (1, 2):
                    _say(f"dither {w}x{h} srccomps={sc} alphabits={ab} dither={d}")
                    out = np.zeros(w * h * 4, np.uint8)
                    L.rgb565_image(out.ctypes.data, src.ctypes.data, w, h, sc, ab, d)
                    exp = oracle.rgb565_image(src, w, h, sc, ab, d)
                    assert np.array_equal(out, exp), (w, h, sc, ab, d)


def test_encode_block_func_per_block(gpu, oracle, clean_env):
    """s2tc_encode_block_func returns per-block entry points with the
    reference's selection rules; block k of a thread uses generator block k."""
    L = gpu.lib()
    FT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int)
    L.s2tc_encode_block_func.restype = ctypes.c_void_p
    L.s2tc_encode_block_func.argtypes = [ctypes.c_int] * 4
    L.s2tc_amd_set_block_index.argtypes = [ctypes.c_uint]
    os.environ["S2TC_RANDOM_SEED"] = "9"
    rng = np.random.default_rng(12)
    W, H = 7, 6
    for dxt in range(3):
        q = imgutil.random_quantized(rng, W, H, dxt, "smooth")
        for cd, nr, refine in [(5, -1, 1), (3, 0, 2), (7, -1, 0), (2, 5, 2), (0, -1, 2), (9, -1, 7)]:
            _say(f"encode_block_func dxt={dxt} cd={cd} nrandom={nr} refine={refine}")
            f = FT(L.s2tc_encode_block_func(dxt, cd, nr, refine))
            bs = 8 if dxt == 0 else 16
            bw, bh = (W + 3) // 4, (H + 3) // 4
            got = np.zeros(bw * bh * bs, np.uint8)
            L.s2tc_amd_set_block_index(0)
            for by in range(bh):
                for bx in range(bw):
                    w, h = min(4, W - 4 * bx), min(4, H - 4 * by)
                    off = ((4 * by) * W + 4 * bx) * 4
                    f(got.ctypes.data + (by * bw + bx) * bs, q.ctypes.data + off, W, w, h, nr)
            exp = oracle.encode_quantized(q, W, H, dxt, cd if cd < 8 else 5, nr, refine if refine < 3 else 1,
                                          seed=9)
            assert np.array_equal(got, exp), (dxt, cd, nr, refine)


def test_concurrent_tx_compress_callers(gpu, clean_env):
    """Calls from several host threads at once give the single-call result."""
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    os.environ.update(S2TC_DITHER_MODE="SIMPLE", S2TC_COLORDIST_MODE="SRGB_MIXED", S2TC_RANDOM_COLORS="0",
                      S2TC_REFINE_COLORS="LOOP")
    rng = np.random.default_rng(13)
    imgs = [imgutil.smooth_rgba8(rng, 96 + 8 * i, 64) for i in range(8)]
    expect = []
    for i, im in enumerate(imgs):
        w = 96 + 8 * i
        dst = np.zeros((w // 4) * 16 * 16, np.uint8)
        L.tx_compress_dxtn(4, w, 64, im.ctypes.data, 0x83F3, dst.ctypes.data, (w // 4) * 16)
        expect.append(dst)
    results = [None] * 8
    errors = []

    def run(i):
        try:
            w = 96 + 8 * i
            for _ in range(3):
                dst = np.zeros((w // 4) * 16 * 16, np.uint8)
                L.tx_compress_dxtn(4, w, 64, imgs[i].ctypes.data, 0x83F3, dst.ctypes.data, (w // 4) * 16)
                if not np.array_equal(dst, expect[i]):
                    errors.append(i)
            results[i] = True
        except Exception as e:  # pragma: no cover
            errors.append(repr(e))
    _say("concurrent tx_compress_dxtn, 8 threads")
    th = [threading.Thread(target=run, args=(i,)) for i in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    assert all(results)


def test_nrandom_limit_is_an_error(gpu):
    """Candidate indices are 15-bit: nrandom above 32752 is refused with an
    error (the output is not written); 32752 itself is accepted."""
    src = gpu.DeviceBuffer(64)
    dst = gpu.DeviceBuffer(16)
    with pytest.raises(gpu.S2tcError, match="nrandom"):
        gpu.encode_image_device(src, 0, 4, 4, dst, 0, 5, 32753, 1)


def test_tx_compress_dxtn_wide_matches_golden(gpu, clean_env):
    """S2TC_RANDOM_COLORS from 209 to 3000 (candidate lists beyond what LDS
    holds: the encoder keeps them in device scratch) through
    tx_compress_dxtn: the reference's bytes for every format."""
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    z = load("wide.npz")
    for key in z["cases"]:
        w, h, fmt, cd, nr, refine = map(int, str(key).split("_"))
        os.environ.update(S2TC_DITHER_MODE="NONE", S2TC_COLORDIST_MODE=NAMES[cd], S2TC_RANDOM_COLORS=str(nr),
                          S2TC_REFINE_COLORS=REFINES[refine], S2TC_RANDOM_SEED="5")
        _say(f"tx_compress_dxtn wide {key}")
        exp = z["dst_" + str(key)]
        dst = np.full(exp.size, 0xAB, np.uint8)
        bs = 8 if fmt < 0x83F2 else 16
        L.tx_compress_dxtn(4, w, h, z[f"src_{w}_{h}"].ctypes.data, fmt, dst.ctypes.data, ((w + 3) // 4) * bs)
        assert np.array_equal(dst, exp), key


@pytest.mark.parametrize("w,h,sc", [(2049, 2049, 3), (2049, 1025, 4), (1531, 1373, 4)])
def test_large_call_ragged_chunks_against_oracle(gpu, oracle, clean_env, w, h, sc):
    """Sources above 8 MiB whose last staging chunk is not a multiple of the
    host threads' 64-byte slices (2049 x 2049 RGB: a last chunk of 4,206,595
    bytes; 2049 x 1025 and 1531 x 1373 RGBA) equal the oracle's
    tx_compress_dxtn byte for byte, the last texels included."""
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    rng = np.random.default_rng(w * 7 + h + sc)
    src = rng.integers(0, 256, w * h * sc).astype(np.uint8)
    bw, bh = (w + 3) // 4, (h + 3) // 4
    for fmt, bs in ((0x83F0, 8), (0x83F3, 16)):
        _say(f"large tx_compress_dxtn {w}x{h} srccomps={sc} format={fmt:#x}, default settings")
        dst = np.zeros(bw * bh * bs, np.uint8)
        L.tx_compress_dxtn(sc, w, h, src.ctypes.data, fmt, dst.ctypes.data, bw * bs)
        r, exp = oracle.compress(sc, w, h, src, fmt, bw * bs, 5, -1, 1, 1)
        assert np.array_equal(dst, exp[:dst.size]), (w, h, sc, fmt)


@pytest.mark.parametrize("w,h,sc,dither,cd,nr,refine", [
    (4096, 1100, 4, 0, 5, -1, 1),    # NONE, RGBA: quantized in the encoder's loads; bands of 1024 rows
    (8192, 1300, 3, 0, 4, -1, 2),    # NONE, RGB: quantized per band; three bands of 512 rows
    (4096, 1401, 4, 1, 1, 8, 2),     # SIMPLE carried over bands, random candidates across bands
    (4096, 1403, 3, 1, 3, 0, 1),     # SIMPLE, RGB, all 16 candidates
])
def test_large_call_bands_against_oracle(gpu, oracle, clean_env, w, h, sc, dither, cd, nr, refine):
    """Calls above one staging chunk and one band of rows (max(512, 2^22 /
    width) rows) are encoded band by band while later rows upload: the SIMPLE
    dither's error state crosses the band boundaries on the device, each
    band's blocks keep their whole-image indices for the random candidates,
    and the last band is ragged (not a multiple of 4 rows); the output equals
    the oracle's tx_compress_dxtn."""
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    os.environ.update(S2TC_DITHER_MODE=DITHERS[dither], S2TC_COLORDIST_MODE=NAMES[cd], S2TC_RANDOM_COLORS=str(nr),
                      S2TC_REFINE_COLORS=REFINES[refine], S2TC_RANDOM_SEED="7")
    rng = np.random.default_rng(w + 3 * h + sc)
    yy, xx = np.mgrid[0:h, 0:w]
    base = np.stack([(xx * 255) // w, (yy * 255) // h, ((xx + yy) * 255) // (w + h), (xx * yy) % 256], -1)
    src = np.clip(base[..., :sc] + rng.integers(-12, 13, (h, w, sc)), 0, 255).astype(np.uint8).reshape(-1)
    assert src.size > 8 << 20
    bw, bh = (w + 3) // 4, (h + 3) // 4
    for fmt, bs in ((0x83F0, 8), (0x83F3, 16)):
        _say(f"banded tx_compress_dxtn {w}x{h} srccomps={sc} dither={DITHERS[dither]} cd={NAMES[cd]} "
             f"nrandom={nr} refine={REFINES[refine]} format={fmt:#x}")
        dst = np.zeros(bw * bh * bs, np.uint8)
        L.tx_compress_dxtn(sc, w, h, src.ctypes.data, fmt, dst.ctypes.data, bw * bs)
        r, exp = oracle.compress(sc, w, h, src, fmt, bw * bs, cd, nr, refine, dither, 7)
        assert r == 0
        assert np.array_equal(dst, exp[:dst.size]), (w, h, sc, dither, cd, nr, fmt)


def test_large_call_staging_and_row_gaps(gpu, clean_env):
    """Sources above the single staging buffer stream through chunked pinned
    buffers, and large outputs are laid out by several host threads: a
    2048x1024 DXT5 call (8 MiB in, 2 MiB of blocks out, 3 MiB with the row
    gap) equals the same call with exact rows placed with the reference's
    dstRowDiff rule, and nothing outside the rows is written."""
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    os.environ.update(S2TC_DITHER_MODE="NONE", S2TC_COLORDIST_MODE="WAVG", S2TC_RANDOM_COLORS="-1",
                      S2TC_REFINE_COLORS="ALWAYS")
    rng = np.random.default_rng(12)
    w, h = 2048, 1024 + 3  # a partial last block row
    src = rng.integers(0, 256, w * h * 4).astype(np.uint8)
    bw, bh, bs = w // 4, (h + 3) // 4, 16
    _say(f"large tx_compress_dxtn {w}x{h}, exact rows")
    exact = np.zeros(bw * bh * bs, np.uint8)
    L.tx_compress_dxtn(4, w, h, src.ctypes.data, 0x83F3, exact.ctypes.data, bw * bs)
    pad = 2048  # dstRowStride - ((w + 3) & ~3) * 4
    stride = bw * bs + pad
    _say(f"large tx_compress_dxtn {w}x{h}, row stride {stride}")
    dst = np.full(stride * bh, 0x5C, np.uint8)
    L.tx_compress_dxtn(4, w, h, src.ctypes.data, 0x83F3, dst.ctypes.data, stride)
    rows = dst.reshape(bh, stride)
    assert np.array_equal(rows[:, :bw * bs].reshape(-1), exact)
    assert (rows[:, bw * bs:] == 0x5C).all()
