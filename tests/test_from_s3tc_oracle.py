"""s2tc_from_s3tc: the oracle's per-index restatement against the reference.

* golden: DDS files of random S3TC blocks and the reference transcoder's
  output (tests/golden/from_s3tc.npz, written by make_golden.py from
  oracle/_ref/s2tc_from_s3tc built from the reference's source);
* live: the same comparison on fresh random files where the reference
  transcoder is built here;
* decoding: every texel of a remapped block decodes (fetch_2d_texel rules)
  to one of the two endpoint colours or, in DXT1's three-colour mode, to
  transparent black, and a second pass decodes like the first;
* the word-wide bit-plane forms the GPU kernel uses, restated in numpy, equal
  the oracle on random blocks.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "from_s3tc.npz"))
FMT = {"DXT1": (0, 8, 0x83F1), "DXT3": (1, 16, 0x83F2), "DXT5": (2, 16, 0x83F3)}


@pytest.mark.parametrize("name", sorted(FMT))
def test_oracle_matches_golden(oracle, name):
    dxt, bs, _ = FMT[name]
    dds, exp = GOLD[f"{name}_in"], GOLD[f"{name}_out"]
    n = (dds.size - 128) // bs
    assert exp.size == 128 + n * bs and np.array_equal(exp[:128], dds[:128])
    got = oracle.from_s3tc(dds[128:128 + n * bs], dxt)
    assert np.array_equal(got, exp[128:])


def test_oracle_matches_live_reference(oracle):
    from oracle import reference_from_s3tc_path
    exe = reference_from_s3tc_path()
    if not exe:
        pytest.skip("reference s2tc_from_s3tc not built here")
    rng = np.random.default_rng(31)
    with tempfile.TemporaryDirectory() as td:
        for name, (dxt, bs, _) in FMT.items():
            blocks = rng.integers(0, 256, (5000, bs)).astype(np.uint8)
            blocks[::3, bs - 6:bs - 4] = blocks[::3, bs - 8:bs - 6]
            hdr = np.zeros(128, np.uint8)
            hdr[84:88] = np.frombuffer(name.encode(), np.uint8)
            i, o = os.path.join(td, "i.dds"), os.path.join(td, "o.dds")
            open(i, "wb").write(hdr.tobytes() + blocks.tobytes())
            subprocess.run([exe, "-i", i, "-o", o], check=True, capture_output=True)
            ref = np.frombuffer(open(o, "rb").read(), np.uint8)[128:]
            assert np.array_equal(oracle.from_s3tc(blocks, dxt), ref), name


def _decode_block(oracle, fmt, blk):
    return np.array([[oracle.fetch(fmt, 4, blk, x, y) for x in range(4)] for y in range(4)])


@pytest.mark.parametrize("name", sorted(FMT))
def test_remapped_blocks_decode_to_endpoints(oracle, name):
    dxt, bs, fmt = FMT[name]
    rng = np.random.default_rng(5 + dxt)
    blocks = rng.integers(0, 256, (300, bs)).astype(np.uint8)
    blocks[::4, bs - 6:bs - 4] = blocks[::4, bs - 8:bs - 6]
    once = oracle.from_s3tc(blocks, dxt).reshape(-1, bs)
    twice = oracle.from_s3tc(once, dxt).reshape(-1, bs)
    for b1, b2 in zip(once, twice):
        d1 = _decode_block(oracle, fmt, b1)
        assert np.array_equal(d1, _decode_block(oracle, fmt, b2))
        cb = b1[bs - 8:]
        c0 = _decode_block(oracle, fmt, np.concatenate([b1[:bs - 8], cb[:4], np.zeros(4, np.uint8)]))[0, 0]
        c1 = _decode_block(oracle, fmt, np.concatenate([b1[:bs - 8], cb[:4], np.full(4, 0x55, np.uint8)]))[0, 0]
        for t in d1.reshape(-1, 4):
            ok = np.array_equal(t[:3], c0[:3]) or np.array_equal(t[:3], c1[:3])
            if dxt == 0 and not ok:
                ok = not t.any()      # transparent black, three-colour mode only
            assert ok, (name, b1, t)
        if dxt == 2:
            a0, a1 = int(b1[0]), int(b1[1])
            assert set(int(v) for v in d1[..., 3].reshape(-1)) <= {a0, a1, 0, 255}
            if a1 < a0:
                assert False, "eight-value alpha mode must not survive the remap"


LO2, CHK2 = 0x55555555, 0x11441144
LO3, CHK3 = 0x249249249249, 0x041208041208


def _swar_colour(w0, w1, dxt1):
    """The kernel's colour remap on (c0 | c1 << 16, index word) arrays."""
    c0, c1 = w0 & 0xFFFF, w0 >> 16
    lo, hi = w1 & LO2, (w1 >> 1) & LO2
    keep3 = (c1 >= c0) & dxt1
    swap = (c1 < c0) if dxt1 else (c1 >= c0)
    r = np.where(keep3, (lo | (hi & CHK2)) | ((hi & lo) << 1), (lo & ~hi & 0xFFFFFFFF) | (hi & CHK2))
    r = np.where(swap, r ^ LO2, r)
    return np.where(swap, c1 | (c0 << 16), w0), r


def _swar_alpha(a):
    a0, a1 = a & 0xFF, (a >> 8) & 0xFF
    idx = a >> 16
    b1, b2 = (idx >> 1) & LO3, (idx >> 2) & LO3
    six = a1 >= a0
    t = np.where(six, b1 ^ b2, b1 | b2)
    r = (idx & ~(t * 7)) | (t & CHK3)
    r = np.where(six, r, r ^ LO3)
    e = np.where(six, a & 0xFFFF, a1 | (a0 << 8))
    return e | (r << 16)


@pytest.mark.parametrize("dxt", [0, 1, 2])
def test_kernel_bitplane_forms(oracle, dxt):
    bs = 8 if dxt == 0 else 16
    rng = np.random.default_rng(77 + dxt)
    blocks = rng.integers(0, 256, (20000, bs)).astype(np.uint8)
    blocks[::9, bs - 6:bs - 4] = blocks[::9, bs - 8:bs - 6]
    exp = oracle.from_s3tc(blocks, dxt).reshape(-1, bs)
    u = blocks.view(np.uint32).astype(np.uint64).reshape(-1, bs // 4)
    cw0, cw1 = _swar_colour(u[:, -2], u[:, -1], dxt == 0)
    got = u.copy()
    got[:, -2], got[:, -1] = cw0, cw1
    if dxt == 2:
        a = _swar_alpha(u[:, 0] | (u[:, 1] << 32))
        got[:, 0], got[:, 1] = a & 0xFFFFFFFF, a >> 32
    assert np.array_equal(got.astype(np.uint32).view(np.uint8).reshape(-1, bs), exp)
