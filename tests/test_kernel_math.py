"""Host checks of the integer identities the GPU kernels use in place of the
reference's direct formulas (exhaustive where the domain is small).

Each identity is restated here in numpy exactly as the kernels compute it and
compared with the reference's formula; a kernel that relied on a false
identity would fail the GPU parity tests too, but these pin the algebra on
the CPU, over whole input domains.
"""
import numpy as np


def test_recip_division_exact():
    """rmean's (2S + n) / (2n) as (x * m) >> 18, m = floor(2^18 / d) + 1, for
    every x < 8192 and d <= 32 (s2tc_device.h, Recip)."""
    x = np.arange(8192, dtype=np.int64)
    for d in range(1, 33):
        m = (1 << 18) // d + 1
        assert np.array_equal((x * m) >> 18, x // d), d
    # the range rmean needs: alpha sums of up to 16 texels, colour sums of 63 per texel
    assert 2 * 16 * 255 + 16 < 8192


def test_alpha_threshold_forms_exhaustive():
    """alpha_try_thr's comparisons against per-pass thresholds equal the
    squared-distance tests of s2tc_try_encode_block (:606-635) for every
    alpha and every pair of endpoints in 0..255."""
    a = np.arange(256, dtype=np.int64)[:, None, None]
    e0 = np.arange(256, dtype=np.int64)[None, :, None]
    e1 = np.arange(256, dtype=np.int64)[None, None, :]
    d0, d1 = (a - e0) ** 2, (a - e1) ** 2
    b_ref = d1 < d0
    bd = np.minimum(d0, d1)
    z_ref = a * a <= bd
    full_ref = ~z_ref & ((255 - a) ** 2 <= bd)

    zh = np.minimum(np.where(e0 == 0, 1023, e0), np.where(e1 == 0, 1023, e1)) >> 1
    fc = (np.maximum(np.where(e0 == 255, 0, 255 + e0), np.where(e1 == 255, 0, 255 + e1)) + 1) >> 1
    s = e0 + e1
    flip = np.where(e1 < e0, 255, 0)
    tb = np.where(e1 > e0, s >> 1, np.where(e1 < e0, 255 - ((s + 1) >> 1), 255))
    z = a <= zh
    full = ~z & (a >= fc)
    b = (a ^ flip) > tb
    assert np.array_equal(z, z_ref)
    assert np.array_equal(full, full_ref)
    acc = ~z_ref & ~full_ref
    # b only matters for texels that go to an endpoint
    assert np.array_equal(b & acc, b_ref & acc)


QUAD = {4: ((4, 1, 4), 1), 5: ((4, 4, 1), 2), 6: ((1, 1, 1), 1)}  # AVG, WAVG, W0AVG: weights, AG


def _dist(w, a, b):
    return sum(w[c] * (a[..., c] - b[..., c]) ** 2 for c in range(3))


def test_quadratic_nearest_endpoint_and_score():
    """encode_quad.h: d1 < d0 <=> sdot4(u, -W, C >> 1) < 0 with int8 weight
    bytes, and the LOOP score rebuilt from block sums equals the sum of
    min(d0, d1), on random and extreme colours."""
    rng = np.random.default_rng(7)
    hi = np.array([31, 63, 31])
    for cd, (w, ag) in QUAD.items():
        for trial in range(300):
            t = rng.integers(0, hi + 1, (16, 3))
            if trial % 3 == 0:
                t = rng.choice([0, 1, 30, 31], (16, 3)) * np.array([1, 2, 1]) + np.array([0, 1, 0]) * rng.integers(0, 2, (16, 3))
                t = np.minimum(t, hi)
            e0, e1 = rng.integers(0, hi + 1, 3), rng.integers(0, hi + 1, 3)
            d0, d1 = _dist(w, t, e0[None]), _dist(w, t, e1[None])
            delta = e1 - e0
            wb = np.array([w[0] * delta[0], (w[1] // ag) * delta[1], w[2] * delta[2]])
            assert np.all(np.abs(wb) <= 127)                     # int8 weight bytes
            u = t * np.array([1, ag, 1])
            assert np.all(u <= 127)                              # non-negative int8 texel bytes
            C = _dist(w, e1, np.zeros(3, int)) - _dist(w, e0, np.zeros(3, int))
            v = (C >> 1) - (u * wb).sum(axis=1)
            assert np.array_equal(v < 0, d1 < d0), (cd, trial)
            # score from sums
            b = d1 < d0
            S = t.sum(axis=0)
            S1 = t[b].sum(axis=0)
            Q = _dist(w, t, np.zeros(3, int)).sum()
            d0sum = Q - 2 * sum(w[c] * S[c] * e0[c] for c in range(3)) + 16 * _dist(w, e0, np.zeros(3, int))
            gain = 2 * sum(w[c] * delta[c] * S1[c] for c in range(3)) - b.sum() * C
            assert d0sum - gain == np.minimum(d0, d1).sum()


def test_first_occurrence_keys():
    """(d << 4 | 15 - k) max and (d << 4 | k) min pick the first texel (in
    slot order k) with the largest / smallest distance, as the reference's
    strict > / < scans do (:900-922)."""
    rng = np.random.default_rng(3)
    for _ in range(2000):
        d = rng.integers(0, rng.choice([3, 20682]), 16)
        kmax = ((d << 4) | (15 - np.arange(16))).max()
        kmin = ((d << 4) | np.arange(16)).min()
        assert 15 - (kmax & 15) == int(np.argmax(d))
        assert (kmin & 15) == int(np.argmin(d))


def test_bytes_to_slots_multiply():
    """bytes_to_slots: 0/1 bytes x = 0..3 -> bit 2x via (m * 0x01041040) >> 24."""
    for v in range(16):
        m = sum(((v >> x) & 1) << (8 * x) for x in range(4))
        got = ((m * 0x01041040) & 0xFFFFFFFF) >> 24
        want = sum(((v >> x) & 1) << (2 * x) for x in range(4))
        assert got == want, v


def test_pair_key_packing_fits():
    """fold_pairs_key: 16 terms of the largest bounded distance plus the 2^23
    penalty stay below 2^24, so (sum << 8 | i << 4 | j) fits 32 bits and real
    sums never reach the penalty."""
    for bound in (400100, 11657, 5891, 20681, 16256):  # NORMALMAP, AVG, W0AVG, WAVG, alpha with fixpoints
        assert 16 * bound < (1 << 23)
        assert (16 * bound + (1 << 23)) << 8 < (1 << 32)


def _uniform_alpha_half(a):
    """alpha_uniform_half of s2tc_device.h: the DXT5 alpha bytes of a full block with one alpha."""
    all6, all7 = sum(6 << (3 * k) for k in range(16)), (1 << 48) - 1
    if a == 0:
        return 0x0100 | (all6 << 16)
    if a == 255:
        return 0xFFFE | (all7 << 16)
    return a | ((a + 1) << 8)


def test_uniform_alpha_closed_form(oracle):
    """A full DXT5 block whose 16 alphas are equal encodes its alpha half as
    alpha_uniform_half says, for every alpha, refinement, quick and
    all-candidate modes and distances (the oracle restates the reference)."""
    rng = np.random.default_rng(12)
    for refine in range(3):
        for cd, nrandom in [(5, -1), (2, -1), (3, -1), (7, -1), (0, 0), (7, 0), (4, 0)]:
            q = np.zeros((4, 4 * 256, 4), np.uint8)
            q[..., 0] = rng.integers(0, 32, (4, 1024))
            q[..., 1] = rng.integers(0, 64, (4, 1024))
            q[..., 2] = rng.integers(0, 32, (4, 1024))
            q[..., 3] = np.repeat(np.arange(256, dtype=np.uint8), 4)[None, :]
            out = oracle.encode_quantized(q.reshape(-1), 1024, 4, 2, cd, nrandom, refine, 0).reshape(-1, 16)
            for a in range(256):
                got = int.from_bytes(out[a, :8].tobytes(), "little")
                assert got == _uniform_alpha_half(a), (refine, cd, nrandom, a)


def test_pair_sums_of_unbounded_distances_never_wrap(oracle):
    """RGB, YUV and SRGB_MIXED are sums of non-negative terms, and their
    16-term pair sums stay below 2^31, so the reference's int sums never wrap
    and its fold "bestsum < 0 || sum < bestsum" is a plain first minimum for
    them (the kernels fold them as non-negative; only SRGB wraps).  Each
    distance is a convex quadratic form g of the feature difference, less at
    most 1 for its rounding shifts, and a convex form takes its largest value
    over the colours between two vertices of the convex hull of their
    (Y, U, V) features: so the largest distance is at most the largest
    vertex-pair distance (from the oracle) + 2."""
    from scipy.spatial import ConvexHull
    r, g, b = (a.ravel() for a in np.meshgrid(np.arange(32), np.arange(64), np.arange(32), indexing="ij"))
    y_srgb = np.floor(np.sqrt((37 * (r * r * 84 + g * g * 72 + b * b * 28)).astype(np.float32))
                      + np.float32(0.5)).astype(np.int64)
    feats = {
        0: (r * 42 + g * 72 + b * 14, 202),   # RGB
        1: (r * 60 + g * 59 + b * 22, 202),   # YUV
        3: (y_srgb, 191),                     # SRGB_MIXED
    }
    for cd, (y, k) in feats.items():
        pts = np.stack([y, r * k - y, b * k - y], 1).astype(float)
        hull = np.unique(ConvexHull(pts).vertices)
        best = 0
        for i in hull:
            for j in hull:
                best = max(best, oracle.color_dist(cd, (int(r[i]), int(g[i]), int(b[i])),
                                                   (int(r[j]), int(g[j]), int(b[j]))))
        assert 16 * (best + 2) < 2 ** 31, (cd, best)


def test_candidate_dedup_equals_full_scan():
    """The random-candidate kernel's colour scan over the list without
    repeated colours (first occurrences, plus each repeated colour's pair
    (first, second) scored by its row sum) picks the same pair as
    reduce_colors_inplace's full scan (:393-410: first pair in (i, j) order
    with the least sum), for symmetric distances with d(c, c) = 0.  Random
    lists with many repeats and small distance ranges, so that ties are
    common."""
    rng = np.random.default_rng(11)
    for trial in range(2000):
        n = int(rng.integers(2, 17))
        m = n + int(rng.integers(1, 65))
        ncol = int(rng.integers(1, 9))
        cols = rng.integers(0, ncol, m)
        # a symmetric distance between colours with a zero diagonal, small values for ties
        dc = rng.integers(0, 4, (ncol, ncol))
        dc = np.triu(dc, 1)
        dc = dc + dc.T
        rows = dc[cols][:, cols[:n]]          # dists[i][k] = d(c_i, c_k), k < n
        # the reference's fold
        best, bi, bj = -1, 0, 1
        for i in range(m):
            for j in range(i + 1, m):
                s = int(np.minimum(rows[i], rows[j]).sum())
                if best < 0 or s < best:
                    best, bi, bj = s, i, j
        # the kernel's: keys (sum, i << 16 | j) over first-occurrence pairs and (first, second) pairs
        first, second = {}, {}
        for i, c in enumerate(cols):
            if c not in first:
                first[c] = i
            elif c not in second:
                second[c] = i
        ul = sorted(first.values())
        keys = []
        for a in range(len(ul)):
            for b in range(a + 1, len(ul)):
                s = int(np.minimum(rows[ul[a]], rows[ul[b]]).sum())
                keys.append((s, ul[a] << 16 | ul[b]))
            c = cols[ul[a]]
            if c in second:
                keys.append((int(rows[ul[a]].sum()), ul[a] << 16 | second[c]))
        s, order = min(keys)
        assert (order >> 16, order & 0xffff) == (bi, bj) and s == best, (trial, n, m, ncol)
