"""Write the golden fixtures under tests/golden/ from the compiled reference.

Run in the container that has /root/reference (oracle/_ref built by
`make -C oracle ref`):  python tests/golden/make_golden.py [--skip-large]

Inputs are seeded synthetic images (s2tc_amd.synth, tests/imgutil) and the
reference's own test images (tests/golden/tga/*.tga, copied from the
reference's tests/).  Expected outputs come only from oracle/_ref, i.e. the
unmodified reference sources with the counter-based rand() replacement.
Files written:
  blocks.npz    block corpus: quantized inputs and expected blocks for every
                (dxt, distance, nrandom, refine)
  dither.npz    rgb565_image outputs for every dither/srccomps/alphabits
  compress.npz  tx_compress_dxtn outputs (environment settings, row strides)
  wide.npz      tx_compress_dxtn outputs with S2TC_RANDOM_COLORS above the
                GPU's LDS-resident limit (209 .. 3000 random colours)
  digests.json  SHA-256 of full-size workload outputs and of the reference
                CLI's DDS files for the test images
  from_s3tc.npz DDS files of random S3TC blocks and the reference
                s2tc_from_s3tc's output for each (oracle/_ref/s2tc_from_s3tc)
  digests.json "decompress": SHA-256 of the reference s2tc_decompress's TGA
                for the reference CLI's DDS of each test image and format
  digests.json "workloads"/"serve_pool": per-texture SHA-256 of the serving
                pool's encodings (every mip level), one per serving request
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

from oracle import Reference  # noqa: E402
import imgutil  # noqa: E402
from s2tc_amd import synth  # noqa: E402
from s2tc_amd import workloads  # noqa: E402

NAMES = ["RGB", "YUV", "SRGB", "SRGB_MIXED", "AVG", "WAVG", "W0AVG", "NORMALMAP"]
REFINES = ["NEVER", "ALWAYS", "LOOP"]
DITHERS = ["NONE", "SIMPLE", "FLOYDSTEINBERG"]
CORPUS_NRANDOM = [-1, 0, 3, 16]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def block_corpus(ref):
    rng = np.random.default_rng(20240101)
    images = {}
    for dxt in range(3):
        lst = []
        for (w, h, kind) in [(13, 9, "random"), (16, 8, "smooth"), (9, 9, "speckle"), (8, 8, "extreme"),
                             (5, 3, "random"), (1, 1, "smooth"), (2, 7, "speckle")]:
            lst.append((w, h, imgutil.random_quantized(rng, w, h, dxt, kind)))
        q, w, h = imgutil.special_blocks(dxt)
        lst.append((w, h, q))
        images[dxt] = lst
    out = {}
    for dxt in range(3):
        for i, (w, h, q) in enumerate(images[dxt]):
            out[f"in_{dxt}_{i}"] = q
            out[f"wh_{dxt}_{i}"] = np.array([w, h], np.int32)
            for cd in range(8):
                for nr in CORPUS_NRANDOM:
                    for refine in range(3):
                        out[f"out_{dxt}_{i}_{cd}_{nr}_{refine}"] = ref.encode_quantized(q, w, h, dxt, cd, nr, refine,
                                                                                      seed=1000 + i)
    np.savez_compressed(os.path.join(HERE, "blocks.npz"), **out)
    return len([k for k in out if k.startswith("out_")])


def dither_corpus(ref):
    rng = np.random.default_rng(7)
    out = {}
    n = 0
    for (w, h) in [(37, 23), (16, 17), (1, 5), (64, 2)]:
        for sc in (3, 4):
            src = rng.integers(0, 256, w * h * sc).astype(np.uint8)
            out[f"src_{w}_{h}_{sc}"] = src
            for ab in (1, 4, 8):
                for d in range(3):
                    out[f"q_{w}_{h}_{sc}_{ab}_{d}"] = ref.rgb565_image(src, w, h, sc, ab, d)
                    n += 1
    np.savez_compressed(os.path.join(HERE, "dither.npz"), **out)
    return n


def compress_corpus(ref):
    """tx_compress_dxtn through the reference with environment settings."""
    rng = np.random.default_rng(11)
    out = {}
    cases = []
    for (w, h) in [(37, 23), (16, 16), (3, 3)]:
        src = rng.integers(0, 256, w * h * 4).astype(np.uint8)
        out[f"src_{w}_{h}"] = src
        for fmt in (0x83F0, 0x83F1, 0x83F2, 0x83F3):
            k = 2 if fmt < 0x83F2 else 4
            bs = 8 if fmt < 0x83F2 else 16
            bw, bh = (w + 3) // 4, (h + 3) // 4
            for stride in (bw * bs, bw * bs + 24, w * k + 1, 0):
                for d in range(3):
                    for (cd, nr, refine) in [(5, -1, 1), (2, 2, 2), (7, 0, 0), (3, -1, 2)]:
                        env = dict(S2TC_DITHER_MODE=DITHERS[d], S2TC_COLORDIST_MODE=NAMES[cd],
                                   S2TC_RANDOM_COLORS=str(nr), S2TC_REFINE_COLORS=REFINES[refine],
                                   S2TC_RANDOM_SEED="77")
                        os.environ.update(env)
                        n = max(stride, bw * bs) * bh + 64
                        dst = np.full(n, 0xAB, np.uint8)
                        ref.lib.tx_compress_dxtn(4, w, h, src.ctypes.data, fmt, dst.ctypes.data, stride)
                        key = f"{w}_{h}_{fmt}_{stride}_{d}_{cd}_{nr}_{refine}"
                        out["dst_" + key] = dst
                        cases.append(key)
    for k in ("S2TC_DITHER_MODE", "S2TC_COLORDIST_MODE", "S2TC_RANDOM_COLORS", "S2TC_REFINE_COLORS",
              "S2TC_RANDOM_SEED"):
        os.environ.pop(k, None)
    out["cases"] = np.array(cases)
    np.savez_compressed(os.path.join(HERE, "compress.npz"), **out)
    return len(cases)


def wide_corpus(ref):
    """tx_compress_dxtn through the reference with many random colours: the
    reference sizes its candidate and distance arrays by 16 + nrandom
    (s2tc_algorithm.cpp:875-876), so any positive value is legal."""
    rng = np.random.default_rng(29)
    out = {}
    cases = []
    for (w, h) in [(13, 9), (8, 8)]:
        src = rng.integers(0, 256, w * h * 4).astype(np.uint8)
        src[3::4][rng.random(w * h) < 0.3] = 0  # some transparent texels for DXT1
        out[f"src_{w}_{h}"] = src
    settings = [(7, 209, 2), (7, 300, 1), (2, 256, 2), (0, 300, 2), (3, 1000, 0), (5, 3000, 1), (1, 433, 2)]
    for (w, h) in [(13, 9), (8, 8)]:
        src = out[f"src_{w}_{h}"]
        bw, bh = (w + 3) // 4, (h + 3) // 4
        for fmt in (0x83F1, 0x83F2, 0x83F3):
            bs = 8 if fmt < 0x83F2 else 16
            for (cd, nr, refine) in settings:
                if nr == 3000 and (w, h) != (8, 8):
                    continue
                env = dict(S2TC_DITHER_MODE="NONE", S2TC_COLORDIST_MODE=NAMES[cd], S2TC_RANDOM_COLORS=str(nr),
                           S2TC_REFINE_COLORS=REFINES[refine], S2TC_RANDOM_SEED="5")
                os.environ.update(env)
                dst = np.full(bw * bh * bs, 0xAB, np.uint8)
                ref.lib.tx_compress_dxtn(4, w, h, src.ctypes.data, fmt, dst.ctypes.data, bw * bs)
                key = f"{w}_{h}_{fmt}_{cd}_{nr}_{refine}"
                out["dst_" + key] = dst
                cases.append(key)
    for k in ("S2TC_DITHER_MODE", "S2TC_COLORDIST_MODE", "S2TC_RANDOM_COLORS", "S2TC_REFINE_COLORS",
              "S2TC_RANDOM_SEED"):
        os.environ.pop(k, None)
    out["cases"] = np.array(cases)
    np.savez_compressed(os.path.join(HERE, "wide.npz"), **out)
    return len(cases)


def cli_digests():
    """DDS files of the reference CLI (oracle/_ref/s2tc_compress with the
    reference library) for the reference's test images and settings."""
    cli = os.path.join(ROOT, "oracle", "_ref", "s2tc_compress")
    lib = os.path.join(ROOT, "oracle", "_ref", "libs2tc_ref.so")
    res = {}
    tga_dir = os.path.join(HERE, "tga")
    for name in sorted(os.listdir(tga_dir)):
        for fmt in ("DXT1", "DXT3", "DXT5"):
            for setting, env in workloads.CLI_SETTINGS.items():
                with tempfile.TemporaryDirectory() as td:
                    o = os.path.join(td, "o.dds")
                    e = dict(os.environ, **env)
                    subprocess.run([cli, "-i", os.path.join(tga_dir, name), "-o", o, "-t", fmt, "-l", lib],
                                   check=True, env=e, capture_output=True)
                    res[f"{name}/{fmt}/{setting}"] = hashlib.sha256(open(o, "rb").read()).hexdigest()
    return res


def cli_workload_digests():
    """The reference CLI (default settings: SIMPLE dither, WAVG, quick mode,
    REFINE_ALWAYS, whole mip chain in the DDS) on the 256x256 sinusoid
    workload written as a 32-bit TGA, DXT1."""
    cli = os.path.join(ROOT, "oracle", "_ref", "s2tc_compress")
    lib = os.path.join(ROOT, "oracle", "_ref", "libs2tc_ref.so")
    wl = workloads.IMAGE_WORKLOADS["sinus256"]
    src = workloads.make_image_source(wl)
    res = {}
    env = {k: v for k, v in os.environ.items() if not k.startswith("S2TC_")}
    with tempfile.TemporaryDirectory() as td:
        t, o = os.path.join(td, "i.tga"), os.path.join(td, "o.dds")
        imgutil.write_tga32(t, src, wl["width"], wl["height"])
        subprocess.run([cli, "-i", t, "-o", o, "-t", "DXT1", "-l", lib], check=True, env=env, capture_output=True)
        res["sinus256"] = {"tga_sha256": hashlib.sha256(open(t, "rb").read()).hexdigest(),
                           "dds_sha256": hashlib.sha256(open(o, "rb").read()).hexdigest()}
    return res


def decompress_digests():
    """TGA files of the reference decoder (oracle/_ref/s2tc_decompress with
    the reference library) for the reference CLI's DDS files of the test
    images, every format, setting faster-wavg-r."""
    cli = os.path.join(ROOT, "oracle", "_ref", "s2tc_compress")
    dec = os.path.join(ROOT, "oracle", "_ref", "s2tc_decompress")
    lib = os.path.join(ROOT, "oracle", "_ref", "libs2tc_ref.so")
    res = {}
    tga_dir = os.path.join(HERE, "tga")
    env = dict(os.environ, **workloads.CLI_SETTINGS["faster-wavg-r"])
    for name in sorted(os.listdir(tga_dir)):
        for fmt in ("DXT1", "DXT3", "DXT5"):
            with tempfile.TemporaryDirectory() as td:
                d, t = os.path.join(td, "o.dds"), os.path.join(td, "o.tga")
                subprocess.run([cli, "-i", os.path.join(tga_dir, name), "-o", d, "-t", fmt, "-l", lib],
                               check=True, env=env, capture_output=True)
                subprocess.run([dec, "-i", d, "-o", t, "-l", lib], check=True, capture_output=True)
                res[f"{name}/{fmt}"] = hashlib.sha256(open(t, "rb").read()).hexdigest()
    return res


def workload_digests(ref, skip_large):
    res = {}
    for name, wl in workloads.IMAGE_WORKLOADS.items():
        if skip_large and wl["width"] * wl["height"] > 1 << 20:
            continue
        src = workloads.make_image_source(wl)
        q = ref.rgb565_image(src, wl["width"], wl["height"], wl["srccomps"], {0: 1, 1: 4, 2: 8}[wl["dxt"]],
                             wl["dither"])
        out, _ = ref.encode_quantized_mt(q, wl["width"], wl["height"], wl["dxt"], wl["cd"], wl["nrandom"],
                                         wl["refine"], seed=wl["seed"])
        res[name] = {"source_sha256": sha(src), "quantized_sha256": sha(q), "blocks_sha256": sha(out)}
        print(name, res[name], flush=True)
    if not skip_large:
        jobs = workloads.batch_jobs(workloads.BATCH_SEED, workloads.BATCH_COUNT)
        digest = batch_digest_reference(ref, jobs)
        res["batch"] = digest
        print("batch", digest, flush=True)
    return res


def reference_texture_blocks(ref, jobs, chunk=400):
    """Reference encode of every (texture, level) job, DITHER_NONE
    quantization; yields (texture index, [blocks of each level])."""
    import ctypes

    class RefJob(ctypes.Structure):
        _fields_ = [("q", ctypes.c_void_p), ("dest", ctypes.c_void_p), ("width", ctypes.c_int),
                    ("height", ctypes.c_int), ("dxt", ctypes.c_int), ("cd", ctypes.c_int),
                    ("nrandom", ctypes.c_int), ("refine", ctypes.c_int), ("seed", ctypes.c_uint64)]
    ref.lib.ref_encode_jobs_mt.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    for t0 in range(0, len(jobs.textures), chunk):
        texs = jobs.textures[t0:t0 + chunk]
        keep, arr, outs = [], [], []
        for tex in texs:
            lv_out = []
            for lv in tex["levels"]:
                w, h = lv["width"], lv["height"]
                q = ref.rgb565_image(lv["rgba"], w, h, 4, {0: 1, 1: 4, 2: 8}[tex["dxt"]], 0)
                dst = np.zeros(((w + 3) // 4) * ((h + 3) // 4) * (8 if tex["dxt"] == 0 else 16), np.uint8)
                keep.append(q)
                lv_out.append(dst)
                arr.append(RefJob(q.ctypes.data, dst.ctypes.data, w, h, tex["dxt"], tex["cd"], tex["nrandom"],
                                  tex["refine"], lv["seed"]))
            outs.append(lv_out)
        ja = (RefJob * len(arr))(*arr)
        ref.lib.ref_encode_jobs_mt(ctypes.cast(ja, ctypes.c_void_p), len(arr), 0)
        for k, lv_out in enumerate(outs):
            yield t0 + k, lv_out
        print(f"textures {t0 + len(texs)}/{len(jobs.textures)}", flush=True)


def batch_digest_reference(ref, jobs):
    """Digest of the batch = SHA-256 over the concatenated per-job block digests."""
    hashes = []
    for _, lv_out in reference_texture_blocks(ref, jobs):
        hashes += [hashlib.sha256(b.tobytes()).digest() for b in lv_out]
    return {"jobs": len(hashes), "texels": jobs.texels, "sha256_of_job_sha256s": hashlib.sha256(b"".join(hashes)).hexdigest()}


def serve_pool_digests(ref):
    """The serving pool (256 textures of the serving mix: 70% quick mode, 30%
    nrandom 64 + LOOP, seed SERVE_SEED) with full mip chains: per texture -
    one serving request - the SHA-256 of its levels' blocks concatenated,
    base level first."""
    jobs = workloads.batch_jobs(workloads.SERVE_SEED, workloads.SERVE_POOL, serve_mix=True)
    per_tex = [None] * len(jobs.textures)
    njobs = 0
    for ti, lv_out in reference_texture_blocks(ref, jobs, chunk=64):
        per_tex[ti] = hashlib.sha256(b"".join(b.tobytes() for b in lv_out)).hexdigest()
        njobs += len(lv_out)
    return {"textures": len(per_tex), "jobs": njobs, "texels": jobs.texels, "texture_sha256": per_tex}


def s3tc_dds(rng, fourcc, nblocks, bs, tail=3):
    """A DDS file: 128-byte header carrying fourcc, random blocks (some with
    equal endpoints), tail bytes of a partial block."""
    blocks = rng.integers(0, 256, (nblocks, bs)).astype(np.uint8)
    blocks[::7, bs - 6:bs - 4] = blocks[::7, bs - 8:bs - 6]      # c1 == c0
    if bs == 16:
        blocks[::5, 1] = blocks[::5, 0]                          # a1 == a0
    hdr = np.zeros(128, np.uint8)
    hdr[0:4] = np.frombuffer(b"DDS ", np.uint8)
    hdr[4:128] = rng.integers(0, 256, 124)
    hdr[84:88] = np.frombuffer(fourcc, np.uint8)
    return np.concatenate([hdr, blocks.reshape(-1), rng.integers(0, 256, tail).astype(np.uint8)])


def from_s3tc_corpus():
    exe = os.path.join(ROOT, "oracle", "_ref", "s2tc_from_s3tc")
    rng = np.random.default_rng(777)
    out = {}
    for fourcc, bs in ((b"DXT1", 8), (b"DXT3", 16), (b"DXT5", 16)):
        dds = s3tc_dds(rng, fourcc, 2048, bs)
        with tempfile.TemporaryDirectory() as td:
            i, o = os.path.join(td, "i.dds"), os.path.join(td, "o.dds")
            open(i, "wb").write(dds.tobytes())
            subprocess.run([exe, "-i", i, "-o", o], check=True, capture_output=True)
            res = np.frombuffer(open(o, "rb").read(), np.uint8)
        out[f"{fourcc.decode()}_in"] = dds
        out[f"{fourcc.decode()}_out"] = res
    np.savez_compressed(os.path.join(HERE, "from_s3tc.npz"), **out)
    return len(out) // 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-large", action="store_true")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    ref = Reference()
    only = set(a.only.split(",")) if a.only else None
    dig_path = os.path.join(HERE, "digests.json")
    digests = json.load(open(dig_path)) if os.path.exists(dig_path) else {}
    if not only or "blocks" in only:
        print("block cases", block_corpus(ref))
    if not only or "dither" in only:
        print("dither cases", dither_corpus(ref))
    if not only or "compress" in only:
        print("compress cases", compress_corpus(ref))
    if not only or "wide" in only:
        print("wide cases", wide_corpus(ref))
    if not only or "cli" in only:
        digests["cli"] = cli_digests()
    if not only or "cli_workloads" in only:
        digests["cli_workloads"] = cli_workload_digests()
    if not only or "decompress" in only:
        digests["decompress"] = decompress_digests()
    if not only or "from_s3tc" in only:
        print("from_s3tc cases", from_s3tc_corpus())
    if not only or "workloads" in only:
        digests.setdefault("workloads", {}).update(workload_digests(ref, a.skip_large))
    if not only or "serve" in only:
        digests.setdefault("workloads", {})["serve_pool"] = serve_pool_digests(ref)
    json.dump(digests, open(dig_path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
