"""The C-ABI library loads, exports every declared entry point, and its
host-side parts (decoders, argument checks) behave like the reference.

No GPU is needed: nothing here launches a kernel.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["txc_dxtn.h", "s2tc_algorithm.h", "s2tc_amd.h"]


def declared_symbols():
    names = []
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^S2TC_API\s+[\w\s\*]*?\b(\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_declared_symbols_cover_reference_abi():
    names = declared_symbols()
    for s in ["tx_compress_dxtn", "fetch_2d_texel_rgb_dxt1", "fetch_2d_texel_rgba_dxt1",
              "fetch_2d_texel_rgba_dxt3", "fetch_2d_texel_rgba_dxt5", "s2tc_encode_block_func", "rgb565_image"]:
        assert s in names


def test_library_exports_every_declared_symbol():
    import s2tc_amd
    L = s2tc_amd.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", s2tc_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared_symbols()) <= exported
    # nothing but the declared C ABI is exported
    assert exported == set(declared_symbols()), sorted(exported - set(declared_symbols()))


def test_enum_values_match_reference():
    import s2tc_amd as s
    assert (s.DXT1, s.DXT3, s.DXT5) == (0, 1, 2)
    assert (s.RGB, s.YUV, s.SRGB, s.SRGB_MIXED, s.AVG, s.WAVG, s.W0AVG, s.NORMALMAP) == tuple(range(8))
    assert (s.REFINE_NEVER, s.REFINE_ALWAYS, s.REFINE_LOOP) == (0, 1, 2)
    assert (s.DITHER_NONE, s.DITHER_SIMPLE, s.DITHER_FLOYDSTEINBERG) == (0, 1, 2)
    txt = open(os.path.join(ROOT, "include", "s2tc_algorithm.h")).read()
    assert "enum DxtMode { DXT1, DXT3, DXT5 }" in txt
    assert "RGB, YUV, SRGB, SRGB_MIXED, AVG, WAVG, W0AVG, NORMALMAP" in txt


FETCH = {0x83F0: "fetch_2d_texel_rgb_dxt1", 0x83F1: "fetch_2d_texel_rgba_dxt1",
         0x83F2: "fetch_2d_texel_rgba_dxt3", 0x83F3: "fetch_2d_texel_rgba_dxt5"}


@pytest.mark.parametrize("fmt", sorted(FETCH))
def test_fetch_decoders_match_oracle(oracle, fmt):
    import s2tc_amd
    fn = getattr(s2tc_amd.lib(), FETCH[fmt])
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(fmt)
    W, H = 13, 10
    bs = 8 if fmt < 0x83F2 else 16
    blob = rng.integers(0, 256, ((W + 3) // 4) * ((H + 3) // 4) * bs).astype(np.uint8)
    # include 6-value-mode DXT5 blocks (a1 >= a0) and DXT1 blocks with c1 >= c0
    blob[::bs] = 3
    for j in range(H):
        for i in range(W):
            t = np.zeros(4, np.uint8)
            fn(W, blob.ctypes.data, i, j, t.ctypes.data)
            assert np.array_equal(t, oracle.fetch(fmt, W, blob, i, j)), (fmt, i, j)


def test_fetch_against_compiled_reference(reference):
    rng = np.random.default_rng(3)
    import s2tc_amd
    for fmt, name in FETCH.items():
        mine = getattr(s2tc_amd.lib(), name)
        theirs = getattr(reference.lib, name)
        for f in (mine, theirs):
            f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        bs = 8 if fmt < 0x83F2 else 16
        blob = rng.integers(0, 256, 16 * bs).astype(np.uint8)
        for j in range(16):
            for i in range(16):
                a, b = np.zeros(4, np.uint8), np.zeros(4, np.uint8)
                mine(16, blob.ctypes.data, i, j, a.ctypes.data)
                theirs(16, blob.ctypes.data, i, j, b.ctypes.data)
                assert np.array_equal(a, b)


def test_bad_format_writes_nothing():
    """destformat outside 0x83F0..0x83F3: message on stderr, dest untouched
    (s2tc_libtxc_dxtn.cpp:232-235).  Runs in a child so stderr can be read."""
    code = f"""
import ctypes, numpy as np, sys
sys.path.insert(0, {ROOT!r})
import s2tc_amd
L = s2tc_amd.lib()
src = np.zeros(8*8*4, np.uint8); dst = np.full(64, 0xAB, np.uint8)
L.tx_compress_dxtn(4, 8, 8, src.ctypes.data_as(ctypes.c_void_p), 0x1234, dst.ctypes.data_as(ctypes.c_void_p), 16)
assert (dst == 0xAB).all()
print("ok")
"""
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "libdxtn: Bad dstFormat 4660 in tx_compress_dxtn" in r.stderr


def test_oracle_decode_image_against_compiled_reference(oracle, reference):
    """The oracle's whole-image decode equals the compiled reference's
    fetch_2d_texel_* texel by texel on a ragged image, with DXT5 blocks in the
    6-value mode, the opaque (254, 255, all indices 7) block and DXT1 blocks
    with c1 >= c0 among random ones."""
    rng = np.random.default_rng(9)
    W, H = 23, 14
    for fmt, name in FETCH.items():
        theirs = getattr(reference.lib, name)
        theirs.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        bs = 8 if fmt < 0x83F2 else 16
        nb = ((W + 3) // 4) * ((H + 3) // 4)
        blob = rng.integers(0, 256, nb * bs).astype(np.uint8).reshape(nb, bs)
        blob[1::3, 0] = 3
        if fmt == 0x83F3:
            blob[2::4, :8] = [254, 255, 255, 255, 255, 255, 255, 255]
        got = oracle.decode_image(blob.reshape(-1), fmt, W, H).reshape(H, W, 4)
        for j in range(H):
            for i in range(W):
                t = np.zeros(4, np.uint8)
                theirs(W, blob.ctypes.data, i, j, t.ctypes.data)
                assert np.array_equal(got[j, i], t), (fmt, i, j)


@pytest.mark.parametrize("w,h", [(0, 5), (7, 0), (0, 0), (-3, 4)])
def test_empty_images_write_nothing(reference, w, h):
    """Width or height of zero (or below): tx_compress_dxtn and rgb565_image
    write nothing and return, as the reference's block loops do (compared
    against the compiled reference); no device work is started."""
    import s2tc_amd
    src = np.arange(64, dtype=np.uint8)
    for lib in (s2tc_amd.lib(), reference.lib):
        f = lib.tx_compress_dxtn
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p,
                      ctypes.c_int]
        for fmt in (0x83F0, 0x83F1, 0x83F2, 0x83F3):
            dst = np.full(64, 0xA5, np.uint8)
            f(4, w, h, src.ctypes.data, fmt, dst.ctypes.data, 16)
            assert (dst == 0xA5).all(), (lib, fmt, w, h)
        g = lib.rgb565_image
        g.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int]
        out = np.full(64, 0x5A, np.uint8)
        g(out.ctypes.data, src.ctypes.data, w, h, 4, 8, 1)
        assert (out == 0x5A).all(), (lib, w, h)
