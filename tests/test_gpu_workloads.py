"""Full-size workloads and the CLI on the GPU against the compiled reference.

The expected SHA-256 digests (tests/golden/digests.json) were computed by
the compiled reference on the same seeded inputs (tests/golden/make_golden.py);
the inputs themselves are regenerated here and checked against their own
recorded digests first.
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DIGESTS = json.load(open(os.path.join(GOLDEN, "digests.json")))
NAMES = ["RGB", "YUV", "SRGB", "SRGB_MIXED", "AVG", "WAVG", "W0AVG", "NORMALMAP"]
REFINES = ["NEVER", "ALWAYS", "LOOP"]
DITHERS = ["NONE", "SIMPLE", "FLOYDSTEINBERG"]
ENV_KEYS = ("S2TC_DITHER_MODE", "S2TC_COLORDIST_MODE", "S2TC_RANDOM_COLORS", "S2TC_REFINE_COLORS",
            "S2TC_RANDOM_SEED")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture
def clean_env():
    saved = {k: os.environ.get(k) for k in ENV_KEYS}
    yield
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("name", ["sinus256", "tiles4096", "normal8192"])
def test_image_workload_digest(gpu, clean_env, name):
    """tx_compress_dxtn on the full-size workload equals the reference bit for bit."""
    from s2tc_amd import workloads
    wl = workloads.IMAGE_WORKLOADS[name]
    exp = DIGESTS["workloads"][name]
    src = workloads.make_image_source(wl)
    assert sha(src) == exp["source_sha256"], "synthetic input differs from the recorded one"
    w, h = wl["width"], wl["height"]
    os.environ.update(S2TC_DITHER_MODE=DITHERS[wl["dither"]], S2TC_COLORDIST_MODE=NAMES[wl["cd"]],
                      S2TC_RANDOM_COLORS=str(wl["nrandom"]), S2TC_REFINE_COLORS=REFINES[wl["refine"]],
                      S2TC_RANDOM_SEED=str(wl["seed"]))
    L = gpu.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    bs = 8 if wl["dxt"] == 0 else 16
    bw, bh = (w + 3) // 4, (h + 3) // 4
    dst = np.zeros(bw * bh * bs, np.uint8)
    fmt = {0: 0x83F1, 1: 0x83F2, 2: 0x83F3}[wl["dxt"]]
    _say(f"workload {name} {w}x{h} through tx_compress_dxtn")
    L.tx_compress_dxtn(wl["srccomps"], w, h, src.ctypes.data, fmt, dst.ctypes.data, bw * bs)
    assert sha(dst) == exp["blocks_sha256"], f"{name}: blocks differ from the reference"


def _run_cli(cli, lib, tga, fmt, env):
    with tempfile.TemporaryDirectory() as td:
        o = os.path.join(td, "o.dds")
        cmd = [cli, "-i", tga, "-o", o, "-t", fmt]
        if lib:
            cmd += ["-l", lib]
        subprocess.run(cmd, check=True, env=dict(os.environ, **env), capture_output=True, timeout=300)
        return hashlib.sha256(open(o, "rb").read()).hexdigest()


def test_cli_on_reference_test_images(gpu):
    """This project's s2tc_compress (GPU library) on the reference's test
    images with the three settings of the reference's tests/test.sh."""
    from s2tc_amd import workloads
    cli = os.path.join(ROOT, "s2tc_amd", "s2tc_compress")
    for key, digest in sorted(DIGESTS["cli"].items()):
        name, fmt, setting = key.split("/")
        _say(f"cli {key}")
        got = _run_cli(cli, None, os.path.join(GOLDEN, "tga", name), fmt, workloads.CLI_SETTINGS[setting])
        assert got == digest, key


def test_cli_sinus256_default_settings(gpu):
    """The 256x256 sinusoid workload through this project's s2tc_compress
    with the library defaults (no S2TC_* variables: SIMPLE dither, WAVG,
    quick mode, REFINE_ALWAYS), DXT1, the whole mip chain in the DDS: the
    reference CLI's file byte for byte.  The reference CLI loading this
    library writes the same file."""
    import imgutil
    from s2tc_amd import workloads
    wl = workloads.IMAGE_WORKLOADS["sinus256"]
    exp = DIGESTS["cli_workloads"]["sinus256"]
    src = workloads.make_image_source(wl)
    env = {k: v for k, v in os.environ.items() if not k.startswith("S2TC_")}
    with tempfile.TemporaryDirectory() as td:
        t, o = os.path.join(td, "i.tga"), os.path.join(td, "o.dds")
        imgutil.write_tga32(t, src, wl["width"], wl["height"])
        assert hashlib.sha256(open(t, "rb").read()).hexdigest() == exp["tga_sha256"]
        clis = [(os.path.join(ROOT, "s2tc_amd", "s2tc_compress"), None)]
        ref_cli = os.path.join(ROOT, "oracle", "_ref", "s2tc_compress")
        if os.path.exists(ref_cli):
            clis.append((ref_cli, os.path.join(ROOT, "s2tc_amd", "libtxc_dxtn.so")))
        for cli, lib in clis:
            _say(f"sinus256 through {os.path.basename(os.path.dirname(cli))}/s2tc_compress, defaults")
            cmd = [cli, "-i", t, "-o", o, "-t", "DXT1"] + (["-l", lib] if lib else [])
            subprocess.run(cmd, check=True, env=env, capture_output=True, timeout=300)
            assert hashlib.sha256(open(o, "rb").read()).hexdigest() == exp["dds_sha256"], cli


def test_reference_cli_loads_this_library(gpu):
    """Drop-in check: the reference's own CLI (built from its sources) loads
    this project's libtxc_dxtn.so with -l and writes the reference's bytes."""
    from s2tc_amd import workloads
    ref_cli = os.path.join(ROOT, "oracle", "_ref", "s2tc_compress")
    if not os.path.exists(ref_cli):
        pytest.skip("reference CLI not built")
    lib = os.path.join(ROOT, "s2tc_amd", "libtxc_dxtn.so")
    for key, digest in sorted(DIGESTS["cli"].items()):
        name, fmt, setting = key.split("/")
        if name not in ("fract001.tga", "supernova.tga"):
            continue
        _say(f"reference cli -l ours {key}")
        got = _run_cli(ref_cli, lib, os.path.join(GOLDEN, "tga", name), fmt, workloads.CLI_SETTINGS[setting])
        assert got == digest, key


def test_decompress_cli_matches_reference_tga(gpu):
    """This project's s2tc_decompress (GPU decode) on the DDS files this
    project's CLI writes for the reference's test images: the TGA bytes the
    reference s2tc_decompress writes for the reference CLI's DDS files."""
    from s2tc_amd import workloads
    cli = os.path.join(ROOT, "s2tc_amd", "s2tc_compress")
    dec = os.path.join(ROOT, "s2tc_amd", "s2tc_decompress")
    env = dict(os.environ, **workloads.CLI_SETTINGS["faster-wavg-r"])
    with tempfile.TemporaryDirectory() as td:
        for key, digest in sorted(DIGESTS["decompress"].items()):
            name, fmt = key.split("/")
            d, t = os.path.join(td, "o.dds"), os.path.join(td, "o.tga")
            _say(f"decompress {key}")
            subprocess.run([cli, "-i", os.path.join(GOLDEN, "tga", name), "-o", d, "-t", fmt], check=True, env=env,
                           capture_output=True, timeout=300)
            subprocess.run([dec, "-i", d, "-o", t], check=True, capture_output=True, timeout=300)
            assert hashlib.sha256(open(t, "rb").read()).hexdigest() == digest, key
