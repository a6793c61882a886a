"""Benchmark and parity workloads (seeded, synthetic; see synth.py).

IMAGE_WORKLOADS are single images encoded like tx_compress_dxtn with the
given settings (dither 1 = SIMPLE, the library default):
  sinus256   256x256 RGBA8 sinusoids + noise, DXT1, WAVG, nrandom -1, ALWAYS
  tiles4096  4096x4096 RGBA8 gradient tiles + radial alpha with speckle,
             DXT5, SRGB_MIXED, nrandom 0, LOOP
  normal8192 8192x8192 RGB8 normal map of 512 Gaussians, DXT1, NORMALMAP,
             nrandom 64, LOOP, generator keyed by (seed 2, block index)
The batch is 2000 textures with their full mip chains (synth.batch_specs),
quantized with DITHER_NONE; level l of a texture with seed s uses generator
seed s2tc_prng_level_seed(s, l).
CLI_SETTINGS are the three encoder settings of the reference's tests/test.sh
(lines 235-245) for colour images.
"""
import numpy as np

from . import synth

IMAGE_WORKLOADS = {
    "sinus256": dict(width=256, height=256, srccomps=4, gen="sinus", seed=0,
                     dxt=0, cd=5, nrandom=-1, refine=1, dither=1),
    "tiles4096": dict(width=4096, height=4096, srccomps=4, gen="tiles", seed=1,
                      dxt=2, cd=3, nrandom=0, refine=2, dither=1),
    "normal8192": dict(width=8192, height=8192, srccomps=3, gen="normalmap", seed=2,
                       dxt=0, cd=7, nrandom=64, refine=2, dither=1),
}

CLI_SETTINGS = {
    "rand32-mrgb-l": dict(S2TC_DITHER_MODE="FLOYDSTEINBERG", S2TC_COLORDIST_MODE="SRGB_MIXED",
                          S2TC_RANDOM_COLORS="32", S2TC_REFINE_COLORS="LOOP", S2TC_RANDOM_SEED="0"),
    "faster-wavg-r": dict(S2TC_DITHER_MODE="SIMPLE", S2TC_COLORDIST_MODE="WAVG",
                          S2TC_RANDOM_COLORS="-1", S2TC_REFINE_COLORS="ALWAYS"),
    "faster-wavg-l": dict(S2TC_DITHER_MODE="SIMPLE", S2TC_COLORDIST_MODE="WAVG",
                          S2TC_RANDOM_COLORS="-1", S2TC_REFINE_COLORS="LOOP"),
}

BATCH_SEED = 3
BATCH_COUNT = 2000
SERVE_SEED = 4
SERVE_POOL = 256

ALPHABITS = {0: 1, 1: 4, 2: 8}
DITHER_NONE, DITHER_SIMPLE = 0, 1


def make_image_source(wl, nthreads=0):
    w, h = wl["width"], wl["height"]
    if wl["gen"] == "sinus":
        return synth.sinus(w, h, wl["seed"], nthreads)
    if wl["gen"] == "tiles":
        return synth.tiles(w, h, wl["seed"], nthreads)
    return synth.normalmap(w, h, wl["seed"], 512, wl["srccomps"], nthreads)


def level_seed(texture_seed, level):
    """s2tc_prng_level_seed of include/s2tc_prng.h."""
    def mix32(x):
        x &= 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x7feb352d) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 0x846ca68b) & 0xFFFFFFFF
        x ^= x >> 16
        return x
    lo = mix32((texture_seed & 0xFFFFFFFF) ^ ((level * 0x85EBCA6B) & 0xFFFFFFFF))
    hi = mix32(((texture_seed >> 32) & 0xFFFFFFFF) ^ lo ^ level)
    return (hi << 32) | lo


class BatchJobs:
    """Textures with mip chains: textures[i] = {dxt, cd, nrandom, refine,
    kind, seed, levels: [{width, height, rgba, seed}]}."""

    def __init__(self, specs, nthreads=0):
        self.textures = []
        self.texels = 0
        for s in specs:
            base = synth.texture(s, nthreads)
            levels = []
            for li, (w, h, lv) in enumerate(synth.mip_chain(base, s.width, s.height)):
                levels.append(dict(width=w, height=h, rgba=lv, seed=level_seed(s.seed, li)))
                self.texels += w * h
            self.textures.append(dict(dxt=s.dxt, cd=s.cd, nrandom=s.nrandom, refine=s.refine, kind=s.kind,
                                      seed=s.seed, levels=levels, width=s.width, height=s.height))

    def jobs(self):
        for ti, t in enumerate(self.textures):
            for li, lv in enumerate(t["levels"]):
                yield ti, li, t, lv


def batch_textures(seed=BATCH_SEED, count=BATCH_COUNT, serve_mix=False, nthreads=0):
    """The base levels only, for DeviceImages.from_textures (mips built on the device)."""
    out = []
    for s in synth.batch_specs(count, seed, serve_mix):
        out.append(dict(rgba=synth.texture(s, nthreads), width=s.width, height=s.height, dxt=s.dxt, cd=s.cd,
                        nrandom=s.nrandom, refine=s.refine, seed=s.seed))
    return out


def batch_jobs(seed=BATCH_SEED, count=BATCH_COUNT, serve_mix=False, nthreads=0):
    return BatchJobs(synth.batch_specs(count, seed, serve_mix), nthreads)
