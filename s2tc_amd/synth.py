"""Deterministic synthetic inputs for the benchmark workloads (ctypes view of
libs2tc_synth.so).  Same seed -> same bytes on every x86-64 host."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libs2tc_synth.so")
_P = ctypes.c_void_p
_I = ctypes.c_int
_U64 = ctypes.c_uint64
_lib = None


class TexSpec(ctypes.Structure):
    _fields_ = [("width", _I), ("height", _I), ("dxt", _I), ("cd", _I), ("nrandom", _I), ("refine", _I),
                ("kind", _I), ("pad", _I), ("seed", _U64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.s2tc_synth_sinus.argtypes = [_P, _I, _I, _U64, _I]
        L.s2tc_synth_tiles.argtypes = [_P, _I, _I, _U64, _I]
        L.s2tc_synth_normalmap.argtypes = [_P, _I, _I, _I, _I, _U64, _I]
        L.s2tc_synth_batch_specs.argtypes = [_P, _I, _U64, _I]
        L.s2tc_synth_texture.argtypes = [_P, ctypes.POINTER(TexSpec), _I]
        L.s2tc_synth_mip_reduce.argtypes = [_P, _P, ctypes.POINTER(_I), ctypes.POINTER(_I)]
        L.s2tc_synth_mip_count.argtypes = [_I, _I]
        _lib = L
    return _lib


def sinus(w, h, seed, nthreads=0):
    out = np.empty(w * h * 4, np.uint8)
    lib().s2tc_synth_sinus(out.ctypes.data, w, h, seed, nthreads)
    return out


def tiles(w, h, seed, nthreads=0):
    out = np.empty(w * h * 4, np.uint8)
    lib().s2tc_synth_tiles(out.ctypes.data, w, h, seed, nthreads)
    return out


def normalmap(w, h, seed, ngauss=512, comps=3, nthreads=0):
    out = np.empty(w * h * comps, np.uint8)
    lib().s2tc_synth_normalmap(out.ctypes.data, comps, w, h, ngauss, seed, nthreads)
    return out


def batch_specs(n, seed, serve_mix=False):
    arr = (TexSpec * n)()
    lib().s2tc_synth_batch_specs(ctypes.cast(arr, _P), n, seed, 1 if serve_mix else 0)
    return list(arr)


def texture(spec, nthreads=0):
    out = np.empty(spec.width * spec.height * 4, np.uint8)
    lib().s2tc_synth_texture(out.ctypes.data, ctypes.byref(spec), nthreads)
    return out


def mip_chain(rgba, w, h):
    """[(w, h, level RGBA8)] from (w, h) down to 1x1 with Image_MipReduce32 semantics."""
    cur = np.array(rgba, dtype=np.uint8, copy=True)
    levels = [(w, h, cur[:w * h * 4].copy())]
    cw, ch = _I(w), _I(h)
    while not (cw.value == 1 and ch.value == 1):
        lib().s2tc_synth_mip_reduce(cur.ctypes.data, cur.ctypes.data, ctypes.byref(cw), ctypes.byref(ch))
        levels.append((cw.value, ch.value, cur[:cw.value * ch.value * 4].copy()))
    return levels


def mip_count(w, h):
    return lib().s2tc_synth_mip_count(w, h)
