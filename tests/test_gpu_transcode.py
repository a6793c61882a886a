"""s2tc_amd_from_s3tc and the s2tc_from_s3tc CLI on the GPU against the
reference transcoder's golden outputs (tests/golden/from_s3tc.npz) and the
oracle (large random block arrays, in place and out of place)."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "from_s3tc.npz"))
CLI = os.path.join(ROOT, "s2tc_amd", "s2tc_from_s3tc")
FMT = {"DXT1": (0, 8), "DXT3": (1, 16), "DXT5": (2, 16)}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", sorted(FMT))
def test_kernel_matches_golden(gpu, name):
    dxt, bs = FMT[name]
    dds, exp = GOLD[f"{name}_in"], GOLD[f"{name}_out"]
    n = (dds.size - 128) // bs
    _say(f"case from_s3tc golden {name} {n} blocks")
    assert np.array_equal(gpu.from_s3tc(dds[128:128 + n * bs], dxt), exp[128:])


@pytest.mark.parametrize("name", sorted(FMT))
def test_kernel_matches_oracle_large(gpu, oracle, name):
    dxt, bs = FMT[name]
    rng = np.random.default_rng(100 + dxt)
    for n in (1, 255, 257, 1 << 20):
        blocks = rng.integers(0, 256, (n, bs)).astype(np.uint8)
        blocks[::11, bs - 6:bs - 4] = blocks[::11, bs - 8:bs - 6]
        _say(f"case from_s3tc {name} n={n}")
        exp = oracle.from_s3tc(blocks, dxt) if n < (1 << 20) else None
        d_in = gpu.DeviceBuffer.from_host(blocks)
        d_out = gpu.DeviceBuffer(blocks.nbytes)
        try:
            gpu.from_s3tc_device(d_in, d_out, n, dxt)
            gpu.synchronize()
            out = d_out.to_host()
            gpu.from_s3tc_device(d_in, d_in, n, dxt)       # in place
            gpu.synchronize()
            assert np.array_equal(d_in.to_host(), out)
        finally:
            d_in.free()
            d_out.free()
        if exp is None:   # spot-check a large array block by block against the oracle
            idx = rng.integers(0, n, 4096)
            exp = oracle.from_s3tc(blocks[idx], dxt).reshape(-1, bs)
            assert np.array_equal(out.reshape(-1, bs)[idx], exp)
        else:
            assert np.array_equal(out, exp)


def test_bad_arguments(gpu):
    L = gpu.lib()
    d = gpu.DeviceBuffer(64)
    try:
        assert L.s2tc_amd_from_s3tc(d.ptr, d.ptr, 2, 3, None) != 0        # not DXT1/3/5
        assert L.s2tc_amd_from_s3tc(d.ptr + 8, d.ptr, 2, 2, None) != 0    # misaligned 16-byte blocks
        assert L.s2tc_amd_from_s3tc(None, None, 0, 2, None) == 0          # nothing to do
    finally:
        d.free()


def test_cli_matches_reference_files(gpu):
    """The CLI writes the reference transcoder's bytes: header copied, whole
    blocks remapped, the trailing partial block dropped; stdin/stdout work."""
    with tempfile.TemporaryDirectory() as td:
        for name in sorted(FMT):
            i, o = os.path.join(td, "i.dds"), os.path.join(td, "o.dds")
            open(i, "wb").write(GOLD[f"{name}_in"].tobytes())
            _say(f"cli from_s3tc {name}")
            subprocess.run([CLI, "-i", i, "-o", o], check=True, capture_output=True, timeout=120)
            assert open(o, "rb").read() == GOLD[f"{name}_out"].tobytes(), name
            r = subprocess.run([CLI], input=GOLD[f"{name}_in"].tobytes(), capture_output=True, check=True,
                               timeout=120)
            assert r.stdout == GOLD[f"{name}_out"].tobytes(), name
        # unsupported fourcc: message, status 1, output created and empty
        bad = bytearray(GOLD["DXT1_in"].tobytes())
        bad[84:88] = b"ATI2"
        open(i, "wb").write(bytes(bad))
        r = subprocess.run([CLI, "-i", i, "-o", o], capture_output=True, timeout=120)
        assert r.returncode == 1 and b"Only DXT1, DXT3, DXT5 are supported!" in r.stderr
        assert os.path.getsize(o) == 0
