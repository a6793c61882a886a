"""s2tc_decompress's container handling on the CPU: with -l naming the
compiled reference library (oracle/_ref/libs2tc_ref.so, built from the
reference's own sources) the tool decodes through that library's
fetch_2d_texel_* functions, exactly as the reference tool does, so the TGA
header, the texel order and the B/R swap are compared byte for byte with the
reference s2tc_decompress on the reference CLI's DDS files and on a file
with an unsupported format.  The GPU decode path is checked by
tests/test_gpu_workloads.py."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
OURS = os.path.join(ROOT, "s2tc_amd", "s2tc_decompress")
TGA_DIR = os.path.join(ROOT, "tests", "golden", "tga")

pytestmark = pytest.mark.skipif(
    not all(os.path.exists(p) for p in (os.path.join(REF, "s2tc_compress"), os.path.join(REF, "s2tc_decompress"),
                                        os.path.join(REF, "libs2tc_ref.so"), OURS)),
    reason="reference tools or this project's s2tc_decompress not built here")


def _run(exe, args):
    return subprocess.run([exe] + args, capture_output=True, timeout=120)


@pytest.mark.parametrize("fmt", ["DXT1", "DXT3", "DXT5"])
def test_container_bytes_equal_reference_tool(fmt):
    lib = os.path.join(REF, "libs2tc_ref.so")
    env = dict(os.environ, S2TC_DITHER_MODE="SIMPLE", S2TC_COLORDIST_MODE="WAVG", S2TC_RANDOM_COLORS="-1",
               S2TC_REFINE_COLORS="ALWAYS")
    with tempfile.TemporaryDirectory() as td:
        for name in ("fract001.tga", "supernova.tga"):
            d = os.path.join(td, "o.dds")
            subprocess.run([os.path.join(REF, "s2tc_compress"), "-i", os.path.join(TGA_DIR, name), "-o", d, "-t", fmt,
                            "-l", lib], check=True, env=env, capture_output=True, timeout=120)
            a, b = os.path.join(td, "ref.tga"), os.path.join(td, "ours.tga")
            assert _run(os.path.join(REF, "s2tc_decompress"), ["-i", d, "-o", a, "-l", lib]).returncode == 0
            assert _run(OURS, ["-i", d, "-o", b, "-l", lib]).returncode == 0
            assert open(a, "rb").read() == open(b, "rb").read(), (name, fmt)


def test_unsupported_format_matches_reference_tool():
    lib = os.path.join(REF, "libs2tc_ref.so")
    with tempfile.TemporaryDirectory() as td:
        d = os.path.join(td, "bad.dds")
        hdr = bytearray(128)
        hdr[12:16] = (4).to_bytes(4, "little")
        hdr[16:20] = (4).to_bytes(4, "little")
        hdr[84:88] = b"ATI2"
        open(d, "wb").write(bytes(hdr) + bytes(16))
        r1 = _run(os.path.join(REF, "s2tc_decompress"), ["-i", d, "-o", os.path.join(td, "a.tga"), "-l", lib])
        r2 = _run(OURS, ["-i", d, "-o", os.path.join(td, "b.tga"), "-l", lib])
        assert r1.returncode == r2.returncode == 1
        assert r1.stderr == r2.stderr
        assert os.path.getsize(os.path.join(td, "a.tga")) == os.path.getsize(os.path.join(td, "b.tga")) == 0
