"""s2tc_amd - MI355X-native S2TC (DXT1/3/5) encoder.

Python view of the C ABI in libtxc_dxtn.so (built in-tree from
s2tc_amd/csrc by __graft_entry__.build()).  The library is the product; this
module only binds it with ctypes.  Importing it never falls back to anything:
if the shared library is missing, ``lib()`` raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtxc_dxtn.so")

# enum values of s2tc_algorithm.h
DXT1, DXT3, DXT5 = 0, 1, 2
RGB, YUV, SRGB, SRGB_MIXED, AVG, WAVG, W0AVG, NORMALMAP = range(8)
REFINE_NEVER, REFINE_ALWAYS, REFINE_LOOP = 0, 1, 2
DITHER_NONE, DITHER_SIMPLE, DITHER_FLOYDSTEINBERG = 0, 1, 2
SRC_QUANTIZED, SRC_RGBA8 = 0, 1
COLORDIST_NAMES = ["RGB", "YUV", "SRGB", "SRGB_MIXED", "AVG", "WAVG", "W0AVG", "NORMALMAP"]
GL_DXT1_RGB, GL_DXT1_RGBA, GL_DXT3, GL_DXT5 = 0x83F0, 0x83F1, 0x83F2, 0x83F3

_P = ctypes.c_void_p
_I = ctypes.c_int
_U64 = ctypes.c_uint64
_SZ = ctypes.c_size_t

_lib = None


class S2tcError(RuntimeError):
    pass


def lib():
    """Load libtxc_dxtn.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise S2tcError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(make -C s2tc_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    L.s2tc_amd_last_error.restype = ctypes.c_char_p
    L.s2tc_amd_malloc.argtypes = [ctypes.POINTER(_P), _SZ]
    L.s2tc_amd_free.argtypes = [_P]
    L.s2tc_amd_memcpy_h2d.argtypes = [_P, _P, _SZ]
    L.s2tc_amd_memcpy_d2h.argtypes = [_P, _P, _SZ]
    L.s2tc_amd_memset.argtypes = [_P, _I, _SZ]
    L.s2tc_amd_encode_image.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I, _U64, _P]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise S2tcError(f"s2tc error {rc}: {lib().s2tc_amd_last_error().decode(errors='replace')}")
    return rc


def block_bytes(dxt):
    return 8 if dxt == DXT1 else 16


def nblocks(width, height):
    return ((width + 3) // 4) * ((height + 3) // 4)


class DeviceBuffer:
    """Device memory owned by the library's allocator."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _P()
        check(lib().s2tc_amd_malloc(ctypes.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        check(lib().s2tc_amd_memcpy_h2d(b.ptr, arr.ctypes.data_as(_P), arr.nbytes))
        return b

    def to_host(self, nbytes=None, dtype=np.uint8):
        n = self.nbytes if nbytes is None else int(nbytes)
        out = np.empty(n, np.uint8)
        check(lib().s2tc_amd_memcpy_d2h(out.ctypes.data_as(_P), self.ptr, n))
        return out.view(dtype)

    def fill(self, value):
        check(lib().s2tc_amd_memset(self.ptr, int(value), self.nbytes))

    def free(self):
        if self.ptr:
            lib().s2tc_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().s2tc_amd_synchronize())


def encode_image_device(d_src, src_kind, width, height, d_dst, dxt, cd, nrandom, refine, seed=0, stream=None):
    """Launch the encoder on device buffers (asynchronous)."""
    check(lib().s2tc_amd_encode_image(d_src.ptr if isinstance(d_src, DeviceBuffer) else d_src, src_kind,
                                      width, height,
                                      d_dst.ptr if isinstance(d_dst, DeviceBuffer) else d_dst,
                                      dxt, cd, nrandom, refine, seed, stream))


def encode_image(texels, width, height, dxt, cd=WAVG, nrandom=-1, refine=REFINE_ALWAYS, seed=0,
                 src_kind=SRC_QUANTIZED):
    """Host convenience: upload 4-byte texels, encode on the GPU, return the blocks."""
    texels = np.ascontiguousarray(texels, dtype=np.uint8).reshape(-1)
    assert texels.size >= width * height * 4
    src = DeviceBuffer.from_host(texels[:width * height * 4])
    nb = nblocks(width, height) * block_bytes(dxt)
    dst = DeviceBuffer(nb)
    encode_image_device(src, src_kind, width, height, dst, dxt, cd, nrandom, refine, seed)
    synchronize()
    out = dst.to_host(nb)
    src.free()
    dst.free()
    return out
