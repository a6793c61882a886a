"""s2tc_amd_encode_image_noalloc (no device allocation: random-candidate
group memory from the library's static scratch array) against
s2tc_amd_encode_image, the entry point it is a variant of."""
import ctypes
import sys

import numpy as np
import pytest

import imgutil

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _encode(gpu, fn, q, w, h, dxt, cd, nr, refine, seed):
    L = gpu.lib()
    f = getattr(L, fn)
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
    f.restype = ctypes.c_int
    src = gpu.DeviceBuffer.from_host(q)
    nbytes = ((w + 3) // 4) * ((h + 3) // 4) * (8 if dxt == 0 else 16)
    dst = gpu.DeviceBuffer(nbytes)
    dst.fill(0xAB)
    gpu.check(f(src.ptr, 0, w, h, dst.ptr, dxt, cd, nr, refine, seed, None))
    gpu.synchronize()
    out = dst.to_host(nbytes)
    src.free()
    dst.free()
    return out


def test_noalloc_equals_encode_image(gpu):
    """Two calls in a row, the second smaller: a 64x64 NORMALMAP DXT1 image
    with 4000 random colours (1.09 MB of group memory per workgroup: 61
    workgroups fit in the static array, so they stride over the 64 tiles),
    then a 13x9 SRGB DXT5 image with 300; then the LDS-resident and quick
    paths through the same entry point."""
    rng = np.random.default_rng(4242)
    cases = [(64, 64, 0, 7, 4000, 2, 17), (13, 9, 2, 2, 300, 1, 3), (21, 11, 1, 5, 64, 2, 8),
             (17, 5, 0, 3, -1, 1, 0)]
    for (w, h, dxt, cd, nr, refine, seed) in cases:
        q = imgutil.random_quantized(rng, w, h, dxt, "random")
        _say(f"noalloc {w}x{h} dxt={dxt} cd={cd} nrandom={nr} refine={refine}")
        exp = _encode(gpu, "s2tc_amd_encode_image", q, w, h, dxt, cd, nr, refine, seed)
        got = _encode(gpu, "s2tc_amd_encode_image_noalloc", q, w, h, dxt, cd, nr, refine, seed)
        assert np.array_equal(got, exp), (w, h, dxt, cd, nr)


def test_noalloc_rejects_bad_arguments(gpu):
    L = gpu.lib()
    L.s2tc_amd_encode_image_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
    src = gpu.DeviceBuffer(64)
    dst = gpu.DeviceBuffer(16)
    assert L.s2tc_amd_encode_image_noalloc(src.ptr, 0, 4, 4, dst.ptr, 0, 5, 32753, 1, 0, None) == -4
    assert L.s2tc_amd_encode_image_noalloc(None, 0, 4, 4, dst.ptr, 0, 5, 300, 1, 0, None) == -1
    src.free()
    dst.free()
