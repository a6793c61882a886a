"""ctypes access to the CPU checkers under oracle/ (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  Two checkers:

* ``Oracle``    - liboracle.so, the C restatement in s2tc_oracle.c;
* ``Reference`` - _ref/libs2tc_ref.so, the unmodified reference sources
  compiled with ref_harness.cpp (present where it was built; it travels with
  the tree as a built file);
* ``StockReference`` - _ref/libs2tc_ref_stock.so, the same build drawing its
  random candidates from libc's rand(), for comparing encoding error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = ctypes.c_void_p
I = ctypes.c_int
U64 = ctypes.c_uint64


def _ptr(a):
    return a.ctypes.data_as(P)


GL_FORMATS = {0: 0x83F1, 1: 0x83F2, 2: 0x83F3}   # DXT1 (RGBA), DXT3, DXT5


def block_bytes(dxt):
    return 8 if dxt == 0 else 16


class _Base:
    def encode_quantized(self, q, width, height, dxt, cd, nrandom, refine, seed=0):
        """Block loop over a quantized (rgb565_image) buffer -> bytes of all blocks."""
        q = np.ascontiguousarray(q, dtype=np.uint8)
        assert q.size >= width * height * 4
        bw, bh = (width + 3) // 4, (height + 3) // 4
        out = np.zeros(bw * bh * block_bytes(dxt), np.uint8)
        self._encode(_ptr(q), width, height, dxt, cd, nrandom, refine, seed, _ptr(out), 0, bh)
        return out


class Oracle(_Base):
    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle)")
        L = self.lib = ctypes.CDLL(path)
        L.oracle_encode_image_quantized.argtypes = [P, I, I, I, I, I, I, U64, P, I, I]
        L.oracle_encode_block.argtypes = [P, P, I, I, I, I, I, I, I, ctypes.c_uint32]
        L.oracle_rgb565_image.argtypes = [P, P, I, I, I, I, I]
        L.oracle_compress_image.argtypes = [I, I, I, P, ctypes.c_uint, P, I, I, I, I, I, U64]
        L.oracle_fetch_texel.argtypes = [ctypes.c_uint, I, P, I, I, P]
        L.oracle_decode_image.argtypes = [ctypes.c_uint, I, I, P, P]
        L.oracle_mip_reduce.argtypes = [P, P, ctypes.POINTER(I), ctypes.POINTER(I)]
        L.oracle_color_dist.argtypes = [I] * 7
        L.oracle_srgb_get_y.argtypes = [I, I, I]
        L.oracle_colordist_error.argtypes = [P, I, I, I, I, P]
        L.oracle_colordist_error.restype = ctypes.c_int64
        L.oracle_from_s3tc_block.argtypes = [P, I]
        self._encode = L.oracle_encode_image_quantized

    def rgb565_image(self, rgba, width, height, srccomps, alphabits, dither):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(width * height * 4, np.uint8)
        self.lib.oracle_rgb565_image(_ptr(out), _ptr(rgba), width, height, srccomps, alphabits, dither)
        return out

    def compress(self, srccomps, width, height, src, fmt, stride, cd, nrandom, refine, dither, seed=0, fill=0):
        src = np.ascontiguousarray(src, dtype=np.uint8)
        bw, bh = (width + 3) // 4, (height + 3) // 4
        bs = 8 if fmt in (0x83F0, 0x83F1) else 16
        n = max(stride, bw * bs) * bh + 64
        out = np.full(n, fill, np.uint8)
        r = self.lib.oracle_compress_image(srccomps, width, height, _ptr(src), fmt, _ptr(out), stride,
                                           cd, nrandom, refine, dither, seed)
        return r, out

    def mip_chain(self, rgba, width, height):
        """[(w, h, RGBA8 level)] down to 1x1 with Image_MipReduce32 semantics."""
        cur = np.ascontiguousarray(rgba, dtype=np.uint8).copy()
        levels = [(width, height, cur[:width * height * 4].copy())]
        w, h = I(width), I(height)
        while not (w.value == 1 and h.value == 1):
            self.lib.oracle_mip_reduce(_ptr(cur), _ptr(cur), ctypes.byref(w), ctypes.byref(h))
            levels.append((w.value, h.value, cur[:w.value * h.value * 4].copy()))
        return levels

    def fetch(self, fmt, row_stride_texels, blocks, i, j):
        t = np.zeros(4, np.uint8)
        self.lib.oracle_fetch_texel(fmt, row_stride_texels, _ptr(blocks), i, j, _ptr(t))
        return t

    def decode_image(self, blocks, fmt, width, height):
        """fetch_2d_texel_* of every texel, row-major RGBA8 (width*height*4 bytes)."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros(width * height * 4, np.uint8)
        self.lib.oracle_decode_image(fmt, width, height, _ptr(blocks), _ptr(out))
        return out

    def color_dist(self, cd, a, b):
        return self.lib.oracle_color_dist(cd, *a, *b)

    def colordist_error(self, q, width, height, dxt, cd, blocks):
        """Sum over texels of ColorDist(texel, decoded colour) (DXT1: alpha-0 texels skipped)."""
        q = np.ascontiguousarray(q, dtype=np.uint8)
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        return int(self.lib.oracle_colordist_error(_ptr(q), width, height, dxt, cd, _ptr(blocks)))


    def from_s3tc(self, blocks, dxt):
        """s2tc_from_s3tc's remap of every block of a DXT1 (dxt 0), DXT3 (1) or DXT5 (2) block array."""
        bs = block_bytes(dxt)
        out = np.array(blocks, dtype=np.uint8, copy=True).reshape(-1)
        n = out.size // bs
        base = out.ctypes.data
        for i in range(n):
            self.lib.oracle_from_s3tc_block(ctypes.c_void_p(base + i * bs), dxt)
        return out


class Reference(_Base):
    """The compiled reference encoder (oracle/_ref), if it was built."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_ref", "libs2tc_ref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = ctypes.CDLL(path)
        L.ref_encode_image_quantized.argtypes = [P, I, I, I, I, I, I, U64, P, I, I]
        L.ref_encode_image_quantized_mt.argtypes = [P, I, I, I, I, I, I, U64, P, I]
        L.ref_encode_image_quantized_mt.restype = I
        L.rgb565_image.argtypes = [P, P, I, I, I, I, I]
        L.tx_compress_dxtn.argtypes = [I, I, I, P, ctypes.c_uint, P, I]
        L.ref_rand_calls.restype = ctypes.c_uint64
        self._encode = L.ref_encode_image_quantized

    def encode_quantized_mt(self, q, width, height, dxt, cd, nrandom, refine, seed=0, nthreads=0):
        q = np.ascontiguousarray(q, dtype=np.uint8)
        bw, bh = (width + 3) // 4, (height + 3) // 4
        out = np.zeros(bw * bh * block_bytes(dxt), np.uint8)
        n = self.lib.ref_encode_image_quantized_mt(_ptr(q), width, height, dxt, cd, nrandom, refine, seed,
                                                   _ptr(out), nthreads)
        return out, n

    def rgb565_image(self, rgba, width, height, srccomps, alphabits, dither):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(width * height * 4, np.uint8)
        self.lib.rgb565_image(_ptr(out), _ptr(rgba), width, height, srccomps, alphabits, dither)
        return out


class StockReference(_Base):
    """The compiled reference with libc's rand() (oracle/_ref/libs2tc_ref_stock.so).
    Single-threaded: the draws then follow the reference's own block order."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_ref", "libs2tc_ref_stock.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = ctypes.CDLL(path)
        L.ref_encode_image_quantized.argtypes = [P, I, I, I, I, I, I, U64, P, I, I]
        L.ref_stock_srand.argtypes = [ctypes.c_uint]
        self._encode = L.ref_encode_image_quantized

    def srand(self, seed=1):
        self.lib.ref_stock_srand(seed)


def reference_from_s3tc_path():
    """The reference's own s2tc_from_s3tc built from its source, or None."""
    p = os.path.join(HERE, "_ref", "s2tc_from_s3tc")
    return p if os.path.exists(p) else None


def stock_reference_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libs2tc_ref_stock.so"))


def reference_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libs2tc_ref.so"))
