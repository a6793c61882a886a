"""Seeded synthetic test inputs (quantized texels and RGBA8 images)."""
import numpy as np

ALPHABITS = {0: 1, 1: 4, 2: 8}


def random_quantized(rng, w, h, dxt, kind="random"):
    """Quantized texels (r5, g6, b5, a) as rgb565_image would produce them."""
    q = np.zeros((h, w, 4), np.uint8)
    ab = ALPHABITS[dxt]
    if kind == "random":
        q[..., 0] = rng.integers(0, 32, (h, w))
        q[..., 1] = rng.integers(0, 64, (h, w))
        q[..., 2] = rng.integers(0, 32, (h, w))
        q[..., 3] = rng.integers(0, 1 << ab, (h, w))
    elif kind == "smooth":
        base = rng.integers(0, 26, 3)
        q[..., 0] = base[0] + rng.integers(0, 6, (h, w))
        q[..., 1] = 2 * base[1] + rng.integers(0, 12, (h, w))
        q[..., 2] = base[2] + rng.integers(0, 6, (h, w))
        q[..., 3] = rng.integers(0, 1 << ab, (h, w))
    elif kind == "speckle":
        q[..., 0] = rng.integers(0, 32, (h, w))
        q[..., 1] = rng.integers(0, 64, (h, w))
        q[..., 2] = rng.integers(0, 32, (h, w))
        q[..., 3] = np.where(rng.random((h, w)) < 0.5, 0, (1 << ab) - 1)
    elif kind == "extreme":   # saturated colours: exercises the SRGB int32 wrap
        q[..., 0] = rng.choice([0, 1, 30, 31], (h, w))
        q[..., 1] = rng.choice([0, 1, 60, 63], (h, w))
        q[..., 2] = rng.choice([0, 1, 30, 31], (h, w))
        q[..., 3] = rng.choice([0, (1 << ab) - 1], (h, w))
    else:
        raise ValueError(kind)
    return q.reshape(-1)


def flat_blocks(rng, w, h, dxt):
    """Quantized texels whose blocks are mostly of one colour among the texels
    that take part in the search: the top half of the block rows has only
    such blocks, the bottom half mixes them with the exceptions (one texel
    off, a lone opaque texel, random blocks)."""
    ab = (1 << ALPHABITS[dxt]) - 1
    bw, bh = (w + 3) // 4, (h + 3) // 4
    q = np.zeros((bh * 4, bw * 4, 4), np.uint8)
    for by in range(bh):
        for bx in range(bw):
            b = q[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]
            col = (rng.integers(0, 32), rng.integers(0, 64), rng.integers(0, 32))
            kind = rng.integers(0, 4) if by < bh // 2 else rng.integers(0, 8)
            b[..., 0], b[..., 1], b[..., 2] = col
            b[..., 3] = ab
            if kind == 1:    # one colour, alpha varies
                b[..., 3] = rng.integers(0, ab + 1, (4, 4))
            elif kind == 2:  # one colour; transparent texels of other colours
                t = rng.random((4, 4)) < 0.4
                b[..., 3] = np.where(t, 0, ab)
                b[..., 0] = np.where(t, rng.integers(0, 32, (4, 4)), col[0])
            elif kind == 3:  # one colour, one alpha below the maximum
                b[..., 3] = rng.integers(0, ab + 1)
            elif kind == 4:  # one texel of another colour
                b[rng.integers(0, 4), rng.integers(0, 4), 1] ^= 1
            elif kind == 5:  # a lone opaque texel
                b[..., 3] = 0
                b[..., 0] = rng.integers(0, 32, (4, 4))
                b[rng.integers(0, 4), rng.integers(0, 4), 3] = ab
            elif kind == 6:  # everything transparent
                b[..., 3] = 0
            elif kind == 7:  # random block
                b[..., 0] = rng.integers(0, 32, (4, 4))
                b[..., 1] = rng.integers(0, 64, (4, 4))
                b[..., 2] = rng.integers(0, 32, (4, 4))
                b[..., 3] = rng.integers(0, ab + 1, (4, 4))
    return np.ascontiguousarray(q[:h, :w]).reshape(-1)


def small_box_blocks(rng, w, h, dxt):
    """Quantized texels whose blocks span small colour boxes (each channel
    within 1, 2, 4 or 8 steps of a base), so that random candidates repeat
    colours; alpha varies within 4 steps (DXT5), and 1/8 of the blocks carry
    transparent texels (DXT1)."""
    ab = (1 << ALPHABITS[dxt]) - 1
    bw, bh = (w + 3) // 4, (h + 3) // 4
    q = np.zeros((bh * 4, bw * 4, 4), np.uint8)
    for by in range(bh):
        for bx in range(bw):
            b = q[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]
            span = int(rng.choice([1, 2, 4, 8]))
            for c, top in enumerate((32, 64, 32)):
                base = int(rng.integers(0, top - span + 1))
                b[..., c] = base + rng.integers(0, span, (4, 4))
            base_a = int(rng.integers(0, max(1, ab - 3) + 1))
            b[..., 3] = np.minimum(ab, base_a + rng.integers(0, 4, (4, 4))) if dxt else ab
            if dxt == 0 and rng.random() < 0.125:
                b[..., 3] = np.where(rng.random((4, 4)) < 0.3, 0, ab)
    return np.ascontiguousarray(q[:h, :w]).reshape(-1)


def special_blocks(dxt):
    """A 4-block-row image of hand-picked 4x4 blocks (returns q, w, h)."""
    ab = (1 << ALPHABITS[dxt]) - 1
    blocks = []
    def blk(fn):
        b = np.zeros((4, 4, 4), np.uint8)
        for y in range(4):
            for x in range(4):
                b[y, x] = fn(x, y)
        blocks.append(b)
    blk(lambda x, y: (31, 63, 31, ab))                   # uniform white
    blk(lambda x, y: (0, 0, 0, ab))                      # uniform black
    blk(lambda x, y: (12, 40, 7, 0))                     # all transparent
    blk(lambda x, y: (5, 10, 20, ab) if (x + y) & 1 else (25, 50, 2, ab))   # two colours
    blk(lambda x, y: (x * 8, y * 16, (x + y) * 3, ab if (x * y) & 1 else 0))
    blk(lambda x, y: (0, 60, 0, ab) if x < 2 else (30, 0, 30, ab))           # SRGB negative distance pair
    blk(lambda x, y: (31, 63, 31, ab) if x == 0 and y == 0 else (31, 63, 30, ab))
    blk(lambda x, y: (7, 7, 7, ab if x == 3 and y == 3 else 0))              # one opaque texel
    blk(lambda x, y: (x * 10, 63 - y * 20, 31 - x * 10, (x * 60 + y * 15) % (ab + 1)))
    blk(lambda x, y: (0, 0, 0, ab) if (x, y) == (2, 1) else (31, 63, 31, ab))
    n = len(blocks)
    img = np.zeros((4, 4 * n, 4), np.uint8)
    for i, b in enumerate(blocks):
        img[:, 4 * i:4 * i + 4] = b
    return img.reshape(-1), 4 * n, 4


def smooth_rgba8(rng, w, h):
    """Natural-looking RGBA8 content (sinusoids + noise)."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    ph = rng.random(4) * 6.28
    img = np.empty((h, w, 4), np.uint8)
    for c in range(4):
        v = 128 + 100 * np.sin(x / 7.0 + ph[c]) * np.cos(y / 9.0 + ph[c]) + rng.integers(-20, 21, (h, w))
        img[..., c] = np.clip(v, 0, 255).astype(np.uint8)
    return img.reshape(-1)


def write_tga32(path, rgba, w, h):
    """Uncompressed 32-bit true-colour TGA (type 2, top-left origin, BGRA order)."""
    px = np.ascontiguousarray(rgba, dtype=np.uint8).reshape(h, w, 4)
    hdr = bytearray(18)
    hdr[2] = 2
    hdr[12:14] = w.to_bytes(2, "little")
    hdr[14:16] = h.to_bytes(2, "little")
    hdr[16] = 32
    hdr[17] = 0x28
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(px[..., [2, 1, 0, 3]].tobytes())


def read_capture(path):
    """[(w, h, srccomps, bytes)] recorded by oracle/_ref/libtxc_capture.so."""
    data = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(data):
        w, h, c = np.frombuffer(data, np.int32, 3, pos)
        pos += 12
        n = int(w) * int(h) * int(c)
        out.append((int(w), int(h), int(c), np.frombuffer(data, np.uint8, n, pos).copy()))
        pos += n
    return out


def corner_blocks(rng, w, h, dxt):
    """Quantized texels at the corners of the colour cube (each channel 0 or
    its maximum), opaque: blocks whose candidate pairs leave large
    distances, so that the least pair sum of a block can exceed 2^28 for the
    unbounded distances (RGB, YUV, SRGB_MIXED)."""
    ab = (1 << ALPHABITS[dxt]) - 1
    q = np.zeros((h, w, 4), np.uint8)
    for c, top in enumerate((31, 63, 31)):
        q[..., c] = top * rng.integers(0, 2, (h, w))
    q[..., 3] = ab
    return q.reshape(-1)


def _mix32(x):
    x &= 0xffffffff
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xffffffff
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xffffffff
    x ^= x >> 16
    return x


def prng_block_key(seed, block):
    """s2tc_prng_block_key of include/s2tc_prng.h."""
    k = _mix32((seed & 0xffffffff) ^ 0x9E3779B9)
    k = _mix32(k ^ (seed >> 32))
    return _mix32(k ^ block)


def prng_draw(bkey, k):
    """s2tc_prng_draw of include/s2tc_prng.h."""
    return _mix32(bkey + k * 0x9E3779B9) & 0x7FFFFFFF


def least_pair_sums(oracle, q, w, h, dxt, cd, nrandom, seed):
    """Per block of an opaque quantized image: the least sum over candidate
    pairs i < j of sum_k min(d(c_i, t_k), d(c_j, t_k)), with the candidate
    list of the random-candidate mode (the block's texels in x-major order,
    then nrandom colours drawn from their box) and the oracle's distance."""
    q = np.asarray(q, np.uint8).reshape(h, w, 4)
    bw, bh = (w + 3) // 4, (h + 3) // 4
    dpc = 4 if dxt == 2 else 3
    out = []
    for by in range(bh):
        for bx in range(bw):
            tex = [tuple(int(v) for v in q[4 * by + y, 4 * bx + x, :3])
                   for x in range(4) for y in range(4) if 4 * bx + x < w and 4 * by + y < h]
            lo = [min(t[c] for t in tex) for c in range(3)]
            ln = [max(t[c] for t in tex) - lo[c] + 1 for c in range(3)]
            key = prng_block_key(seed, by * bw + bx)
            cand = list(tex)
            for c in range(nrandom):
                cand.append(tuple(lo[z] + prng_draw(key, c * dpc + z) % ln[z] for z in range(3)))
            d = np.array([[oracle.color_dist(cd, ci, tk) for tk in tex] for ci in cand], np.int64)
            best = min(int(np.minimum(d[i], d[i + 1:]).sum(1).min()) for i in range(len(cand) - 1))
            out.append(best)
    return out
