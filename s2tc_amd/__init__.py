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
    L.s2tc_amd_decode_image.argtypes = [_P, ctypes.c_uint, _I, _I, _P, _P]
    L.s2tc_amd_colordist_error.argtypes = [_P, _I, _I, _I, _P, _I, _I, ctypes.POINTER(ctypes.c_longlong), _P]
    L.s2tc_amd_mip_layout.argtypes = [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_SZ), ctypes.POINTER(_I),
                                      ctypes.POINTER(_I), _I, ctypes.POINTER(_SZ)]
    L.s2tc_amd_mip_reduce.argtypes = [_P, _I, _I, _P, _P]
    L.s2tc_amd_mip_chain.argtypes = [_P, _I, _I, _P, _SZ, _P]
    L.s2tc_amd_from_s3tc.argtypes = [_P, _P, _SZ, _I, _P]
    L.s2tc_amd_mip_chain_batch.argtypes = [_P, _I, _P]
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


def _dptr(b):
    return b.ptr if isinstance(b, DeviceBuffer) else b


def rgb565_image(rgba, width, height, srccomps, alphabits, dither):
    """rgb565_image of the C ABI (quantization and dither on the GPU): 8-bit
    RGB(A) host bytes -> width*height quantized texels {r5, g6, b5, a}."""
    L = lib()
    L.rgb565_image.argtypes = [_P, _P, _I, _I, _I, _I, _I]
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8).reshape(-1)
    assert rgba.size >= width * height * (3 if srccomps == 3 else 4)
    out = np.zeros(width * height * 4, np.uint8)
    L.rgb565_image(out.ctypes.data_as(_P), rgba.ctypes.data_as(_P), width, height, srccomps, alphabits, dither)
    return out


def decode_image_device(d_blocks, fmt, width, height, d_rgba, stream=None):
    """fetch_2d_texel_* of every texel into d_rgba (width*height RGBA8), asynchronous."""
    check(lib().s2tc_amd_decode_image(_dptr(d_blocks), fmt, width, height, _dptr(d_rgba), stream))


def decode_image(blocks, fmt, width, height):
    """Host convenience: decode a block image on the GPU -> RGBA8 bytes."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1)
    d_b = DeviceBuffer.from_host(blocks)
    d_o = DeviceBuffer(width * height * 4)
    decode_image_device(d_b, fmt, width, height, d_o)
    synchronize()
    out = d_o.to_host(width * height * 4)
    d_b.free()
    d_o.free()
    return out


def colordist_error_device(d_src, src_kind, width, height, d_blocks, dxt, cd, stream=None):
    """Sum over texels of ColorDist(texel, decoded colour) (synchronous)."""
    total = ctypes.c_longlong(0)
    check(lib().s2tc_amd_colordist_error(_dptr(d_src), src_kind, width, height, _dptr(d_blocks), dxt, cd,
                                         ctypes.byref(total), stream))
    return total.value


def colordist_error(texels, width, height, blocks, dxt, cd, src_kind=SRC_QUANTIZED):
    """Host convenience for colordist_error_device."""
    texels = np.ascontiguousarray(texels, dtype=np.uint8).reshape(-1)
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1)
    d_s, d_b = DeviceBuffer.from_host(texels[:width * height * 4]), DeviceBuffer.from_host(blocks)
    try:
        return colordist_error_device(d_s, src_kind, width, height, d_b, dxt, cd)
    finally:
        d_s.free()
        d_b.free()


# ---- mip chains ----

def mip_layout(width, height):
    """[(offset, w, h)] of every level of a device mip chain buffer, and its size in bytes."""
    L = lib()
    n, total = _I(0), _SZ(0)
    check(L.s2tc_amd_mip_layout(width, height, ctypes.byref(n), None, None, None, 0, ctypes.byref(total)))
    offs, ws, hs = (_SZ * n.value)(), (_I * n.value)(), (_I * n.value)()
    check(L.s2tc_amd_mip_layout(width, height, ctypes.byref(n), offs, ws, hs, n.value, ctypes.byref(total)))
    return [(offs[k], ws[k], hs[k]) for k in range(n.value)], total.value


def mip_reduce_device(d_in, width, height, d_out, stream=None):
    """One Image_MipReduce32 halving on the device (asynchronous)."""
    check(lib().s2tc_amd_mip_reduce(_dptr(d_in), width, height, _dptr(d_out), stream))


def mip_chain_device(d_src, width, height, d_levels, capacity, stream=None):
    """Every level (level 0 = the image) into d_levels as mip_layout places them; returns the level count."""
    n = lib().s2tc_amd_mip_chain(_dptr(d_src), width, height, _dptr(d_levels), capacity, stream)
    if n <= 0:
        check(n if n < 0 else -1)
    return n


class MipJob(ctypes.Structure):
    _fields_ = [("levels", _P), ("width", _I), ("height", _I)]


def mip_chain_batch(jobs, stream=None):
    """Chains of many images in place, one launch per level: jobs is a list of
    (device address of a chain buffer with level 0 filled, width, height).
    Returns after the chains are built."""
    arr = (MipJob * len(jobs))(*[MipJob(p, w, h) for p, w, h in jobs])
    check(lib().s2tc_amd_mip_chain_batch(ctypes.cast(arr, _P), len(jobs), stream))


def mip_chain(rgba, width, height):
    """Host convenience: [(w, h, RGBA8 bytes)] of the chain computed on the GPU."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8).reshape(-1)[:width * height * 4]
    layout, total = mip_layout(width, height)
    d_s, d_l = DeviceBuffer.from_host(rgba), DeviceBuffer(total)
    try:
        mip_chain_device(d_s, width, height, d_l, total)
        synchronize()
        buf = d_l.to_host(total)
    finally:
        d_s.free()
        d_l.free()
    return [(w, h, buf[o:o + w * h * 4].copy()) for o, w, h in layout]


# ---- S3TC -> S2TC transcoding ----

def from_s3tc_device(d_in, d_out, nblocks, dxt, stream=None):
    """s2tc_from_s3tc's block remap of nblocks blocks on the device (asynchronous; in place allowed)."""
    check(lib().s2tc_amd_from_s3tc(_dptr(d_in), _dptr(d_out), nblocks, dxt, stream))


def from_s3tc(blocks, dxt):
    """Host convenience: remapped copy of a DXT1 (dxt 0), DXT3 (1) or DXT5 (2) block array."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1)
    n = blocks.size // block_bytes(dxt)
    if n == 0:
        return blocks[:0].copy()
    d = DeviceBuffer.from_host(blocks[:n * block_bytes(dxt)])
    try:
        from_s3tc_device(d, d, n, dxt)
        synchronize()
        return d.to_host(n * block_bytes(dxt))
    finally:
        d.free()


# ---- batches ----

class Job(ctypes.Structure):
    """s2tc_amd_job"""
    _fields_ = [("src", _P), ("dst", _P), ("width", _I), ("height", _I), ("src_kind", _I), ("dxt", _I),
                ("cd", _I), ("nrandom", _I), ("refine", _I), ("seed", _U64)]


class ClassTime(ctypes.Structure):
    """s2tc_amd_class_time"""
    _fields_ = [("normal", _I), ("cd", _I), ("dxt", _I), ("refine", _I), ("njobs", _I), ("blocks", ctypes.c_longlong),
                ("texels", ctypes.c_longlong), ("bytes", ctypes.c_longlong), ("ms", ctypes.c_double),
                ("form", _I)]


# kernel forms of a batch class (s2tc_amd.h S2TC_FORM_*)
FORM_LANE, FORM_LANE_2WAVES, FORM_ROW16 = 0, 1, 2


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return L
    L.s2tc_amd_batch_create.argtypes = [_P, _I, ctypes.POINTER(_P)]
    L.s2tc_amd_batch_set_schedule.argtypes = [_P, _I, _I]
    L.s2tc_amd_batch_get_schedule.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]
    L.s2tc_amd_batch_autotune.argtypes = [_P, _P, _I, ctypes.POINTER(ctypes.c_double)]
    L.s2tc_amd_batch_set_form.argtypes = [_P, _I, _I]
    L.s2tc_amd_batch_launch.argtypes = [_P, _P]
    L.s2tc_amd_batch_class_times.argtypes = [_P, _P, _P, _I]
    L.s2tc_amd_batch_destroy.argtypes = [_P]
    L.s2tc_amd_encode_jobs_plain.argtypes = [_P, _I, _P]
    L._batch_bound = True
    return L


class DeviceImages:
    """A set of images packed into one device buffer, with one output buffer
    for all their blocks.  items: iterable of dicts with rgba (uint8 array of
    width*height*4), width, height, dxt, cd, nrandom, refine, seed."""

    ALIGN = 256

    def __init__(self, items, src_kind=SRC_RGBA8):
        items = list(items)
        so, s_total = [], 0
        for it in items:
            so.append(s_total)
            s_total += (it["width"] * it["height"] * 4 + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        self._layout(items, so, s_total, src_kind)
        staging = np.zeros(s_total, np.uint8)
        for it, o in zip(items, so):
            a = np.ascontiguousarray(it["rgba"], dtype=np.uint8).reshape(-1)
            staging[o:o + a.size] = a
        check(lib().s2tc_amd_memcpy_h2d(self.src.ptr, staging.ctypes.data_as(_P), s_total))
        del staging
        self.uploaded_bytes = s_total

    def _layout(self, items, so, s_total, src_kind):
        """Allocates the source (s_total bytes, image i at so[i]) and output buffers and the jobs."""
        self.items = items
        do, d_total = [], 0
        for it in items:
            do.append(d_total)
            nb = nblocks(it["width"], it["height"]) * block_bytes(it["dxt"])
            d_total += (nb + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        self.src = DeviceBuffer(max(s_total, 1))
        self.dst = DeviceBuffer(max(d_total, 1))
        self.src_off, self.dst_off = so, do
        self.texels = sum(it["width"] * it["height"] for it in items)
        self.bytes_in = self.texels * 4
        self.bytes_out = sum(nblocks(it["width"], it["height"]) * block_bytes(it["dxt"]) for it in items)
        self.jobs = (Job * len(items))()
        for i, it in enumerate(items):
            self.jobs[i] = Job(self.src.ptr + so[i], self.dst.ptr + do[i], it["width"], it["height"], src_kind,
                               it["dxt"], it["cd"], it["nrandom"], it["refine"], it["seed"])

    @classmethod
    def from_textures(cls, textures, src_kind=SRC_RGBA8):
        """Textures with their full mip chains built on the device
        (Image_MipReduce32, s2tc_amd_mip_chain): only the base levels are
        uploaded.  textures: dicts with rgba (base level), width, height, dxt,
        cd, nrandom, refine, seed; level k of a texture encodes with seed
        s2tc_prng_level_seed(seed, k).  items / jobs list every level, texture
        by texture, base level first."""
        from .workloads import level_seed
        self = cls.__new__(cls)
        textures = list(textures)
        chains, s_total = [], 0
        for t in textures:
            layout, total = mip_layout(t["width"], t["height"])
            chains.append((s_total, layout))
            s_total += (total + cls.ALIGN - 1) // cls.ALIGN * cls.ALIGN
        items, so = [], []
        for t, (base, layout) in zip(textures, chains):
            for k, (off, w, h) in enumerate(layout):
                items.append(dict(rgba=None, width=w, height=h, dxt=t["dxt"], cd=t["cd"], nrandom=t["nrandom"],
                                  refine=t["refine"], seed=level_seed(t["seed"], k)))
                so.append(base + off)
        self._layout(items, so, s_total, src_kind)
        staging = np.zeros(s_total, np.uint8)
        for t, (base, _) in zip(textures, chains):
            a = np.ascontiguousarray(t["rgba"], dtype=np.uint8).reshape(-1)[:t["width"] * t["height"] * 4]
            staging[base:base + a.size] = a
        check(lib().s2tc_amd_memcpy_h2d(self.src.ptr, staging.ctypes.data_as(_P), s_total))
        del staging
        mip_chain_batch([(self.src.ptr + base, t["width"], t["height"]) for t, (base, _) in zip(textures, chains)])
        self.uploaded_bytes = sum(t["width"] * t["height"] * 4 for t in textures)
        return self

    def source(self, i):
        """Host copy of image i's RGBA8 source as it lies on the device."""
        it = self.items[i]
        n = it["width"] * it["height"] * 4
        out = np.empty(n, np.uint8)
        check(lib().s2tc_amd_memcpy_d2h(out.ctypes.data_as(_P), self.src.ptr + self.src_off[i], n))
        return out

    def output(self, i, host=None):
        """Blocks of image i (copies the whole output buffer once if host is None)."""
        it = self.items[i]
        n = nblocks(it["width"], it["height"]) * block_bytes(it["dxt"])
        host = self.dst.to_host() if host is None else host
        return host[self.dst_off[i]:self.dst_off[i] + n]

    def free(self):
        self.src.free()
        self.dst.free()


class Batch:
    """s2tc_amd_batch: the images of a DeviceImages in one graph launch."""

    def __init__(self, images):
        L = _bind_batch(lib())
        self.images = images
        h = _P()
        check(L.s2tc_amd_batch_create(ctypes.cast(images.jobs, _P), len(images.jobs), ctypes.byref(h)))
        self.h = h.value

    def launch(self, stream=None):
        check(lib().s2tc_amd_batch_launch(self.h, stream))

    def set_schedule(self, wg_per_cu, streams):
        check(lib().s2tc_amd_batch_set_schedule(self.h, wg_per_cu, streams))

    def schedule(self):
        a, b = _I(), _I()
        check(lib().s2tc_amd_batch_get_schedule(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, reps=3, stream=None):
        ms = ctypes.c_double()
        check(lib().s2tc_amd_batch_autotune(self.h, stream, reps, ctypes.byref(ms)))
        return ms.value

    def set_form(self, klass, form):
        """Kernel form of class `klass` (index into class_times())."""
        check(lib().s2tc_amd_batch_set_form(self.h, klass, form))

    def class_times(self, stream=None):
        arr = (ClassTime * 64)()
        n = lib().s2tc_amd_batch_class_times(self.h, stream, ctypes.cast(arr, _P), 64)
        if n < 0:
            check(n)
        return [dict(normal=c.normal, cd=c.cd, dxt=c.dxt, refine=c.refine, njobs=c.njobs, blocks=c.blocks,
                     texels=c.texels,
                     bytes=c.bytes, ms=c.ms, form=c.form) for c in arr[:n]]

    def destroy(self):
        if self.h:
            lib().s2tc_amd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def encode_jobs_plain(images, stream=None):
    L = _bind_batch(lib())
    check(L.s2tc_amd_encode_jobs_plain(ctypes.cast(images.jobs, _P), len(images.jobs), stream))


# ---- serving ----

class Request(ctypes.Structure):
    """s2tc_amd_request"""
    _fields_ = [("arrival_s", ctypes.c_double), ("first_job", _I), ("njobs", _I)]


class ServeOpts(ctypes.Structure):
    """s2tc_amd_serve_opts"""
    _fields_ = [("hbm_budget", _SZ), ("streams", _I), ("max_batch", _I), ("serial_plain", _I),
                ("wg_per_cu", _I), ("keep_outputs", ctypes.POINTER(_P))]


class ServeStats(ctypes.Structure):
    """s2tc_amd_serve_stats"""
    _fields_ = [("completed", _I), ("rejected", _I), ("batches", _I), ("max_queue", _I), ("max_inflight", _I),
                ("wall_s", ctypes.c_double), ("texels", ctypes.c_longlong), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double)]


def serve(images, requests, hbm_budget=64 << 30, streams=4, max_batch=256, serial_plain=False, wg_per_cu=16,
          keep_outputs=None):
    """Run requests [(arrival_s, first_job, njobs)] over the images of a
    DeviceImages in real time.  Returns (stats dict, latencies_ms array).
    keep_outputs: optional list (one per image) of uint8 arrays to receive blocks."""
    L = lib()
    L.s2tc_amd_serve.argtypes = [_P, _I, _P, _I, ctypes.POINTER(ServeOpts), _P, ctypes.POINTER(ServeStats)]
    reqs = (Request * len(requests))(*[Request(a, f, n) for (a, f, n) in requests])
    keep = None
    if keep_outputs is not None:
        keep = (_P * len(images.jobs))(*[o.ctypes.data if o is not None else None for o in keep_outputs])
    opts = ServeOpts(int(hbm_budget), streams, max_batch, 1 if serial_plain else 0, wg_per_cu,
                     ctypes.cast(keep, ctypes.POINTER(_P)) if keep is not None else None)
    lat = np.zeros(len(requests), np.float64)
    st = ServeStats()
    check(L.s2tc_amd_serve(ctypes.cast(images.jobs, _P), len(images.jobs), ctypes.cast(reqs, _P), len(requests),
                           ctypes.byref(opts), lat.ctypes.data_as(_P), ctypes.byref(st)))
    stats = {k: getattr(st, k) for k, _ in ServeStats._fields_}
    return stats, lat
