"""The device mip chain (s2tc_amd_mip_chain / s2tc_amd_mip_reduce) against
the oracle's Image_MipReduce32 restatement, which the CPU suite pins to the
reference CLI's own levels: bit-exact on odd, one-texel-wide and
one-texel-high images, a large image, and the argument errors."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (7, 1), (5, 3), (13, 2), (2, 13), (33, 17), (64, 1), (3, 64), (31, 33),
         (256, 256), (1000, 37), (37, 1000)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cmp(gpu, oracle, rgba, w, h):
    got = gpu.mip_chain(rgba, w, h)
    exp = oracle.mip_chain(rgba, w, h)
    assert [(a, b) for a, b, _ in got] == [(a, b) for a, b, _ in exp]
    for k, ((lw, lh, a), (_, _, b)) in enumerate(zip(got, exp)):
        assert np.array_equal(a, b), (w, h, k, lw, lh)


def test_mip_chain_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(21)
    for w, h in SIZES:
        _say(f"case mip_chain {w}x{h}")
        _cmp(gpu, oracle, rng.integers(0, 256, w * h * 4).astype(np.uint8), w, h)


def test_mip_chain_extremes_and_large(gpu, oracle):
    """All-255 and all-0 neighbours (no carry between channels) and a
    4096x2048 image, where every level is compared."""
    rng = np.random.default_rng(22)
    for w, h in [(9, 7), (4096, 2048)]:
        rgba = rng.choice(np.array([0, 1, 254, 255], np.uint8), w * h * 4)
        _say(f"case mip_chain extremes {w}x{h}")
        _cmp(gpu, oracle, rgba, w, h)


def test_mip_layout_and_errors(gpu):
    layout, total = gpu.mip_layout(5, 3)
    assert [(w, h) for _, w, h in layout] == [(5, 3), (2, 1), (1, 1)]
    assert [o for o, _, _ in layout] == [0, 256, 512] and total == 768
    L = gpu.lib()
    d = gpu.DeviceBuffer(1024)
    try:
        assert L.s2tc_amd_mip_reduce(d.ptr, 1, 1, d.ptr, None) != 0      # nothing to reduce
        assert L.s2tc_amd_mip_reduce(d.ptr, 0, 4, d.ptr, None) != 0
        assert L.s2tc_amd_mip_chain(d.ptr, 5, 3, d.ptr, 767, None) != 0  # buffer too small
        src = gpu.DeviceBuffer.from_host(np.arange(60, dtype=np.uint8))
        assert gpu.mip_chain_device(src, 5, 3, d, 1024) == 3
        src.free()
    finally:
        d.free()


def test_mip_chain_batch_matches_oracle(gpu, oracle):
    """s2tc_amd_mip_chain_batch: many chains of mixed sizes, one launch per
    level, each level equal to the oracle's."""
    rng = np.random.default_rng(23)
    sizes = SIZES + [(1, 1), (2, 1), (512, 8), (8, 512), (300, 299)]
    bufs = []
    try:
        for w, h in sizes:
            layout, total = gpu.mip_layout(w, h)
            rgba = rng.integers(0, 256, w * h * 4).astype(np.uint8)
            d = gpu.DeviceBuffer(total)
            gpu.check(gpu.lib().s2tc_amd_memcpy_h2d(d.ptr, rgba.ctypes.data, rgba.nbytes))
            bufs.append((d, w, h, rgba, layout, total))
        _say(f"case mip_chain_batch {len(bufs)} chains")
        gpu.mip_chain_batch([(d.ptr, w, h) for d, w, h, _, _, _ in bufs])
        for d, w, h, rgba, layout, total in bufs:
            host = d.to_host(total)
            exp = oracle.mip_chain(rgba, w, h)
            assert len(exp) == len(layout)
            for (o, lw, lh), (ew, eh, lv) in zip(layout, exp):
                assert (lw, lh) == (ew, eh)
                assert np.array_equal(host[o:o + lw * lh * 4], lv), (w, h, lw, lh)
    finally:
        for b in bufs:
            b[0].free()
    assert gpu.lib().s2tc_amd_mip_chain_batch(None, 0, None) == 0
