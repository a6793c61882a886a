"""The serving loop: every request's blocks equal the compiled reference's
(per-texture digests of the serving pool, tests/golden/digests.json) under
Poisson arrivals, are the same alone and in the mix, and budget admission
queues or rejects instead of failing."""
import hashlib
import json
import os
import sys
import types

import numpy as np
import pytest

from test_gpu_batch import small_batch_items

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _texture_requests(items):
    """Group consecutive images of the same texture (mip chains) into requests."""
    groups, start = [], 0
    for i in range(1, len(items) + 1):
        if i == len(items) or (items[i]["width"] >= items[i - 1]["width"] and items[i]["height"] >= items[i - 1]["height"]):
            groups.append((start, i - start))
            start = i
    return groups


def test_serve_results_equal_alone_and_in_mix(gpu):
    items = small_batch_items(21, 40, max_k=8, serve_mix=True)
    imgs = gpu.DeviceImages(items)
    groups = _texture_requests(items)
    # expected: each image encoded alone with a plain launch
    imgs.dst.fill(0)
    gpu.encode_jobs_plain(imgs)
    gpu.synchronize()
    host = imgs.dst.to_host()
    alone = [imgs.output(i, host).copy() for i in range(len(items))]
    rng = np.random.default_rng(2)
    # 3 rounds of the same textures arriving in a burst: they batch together
    reqs = []
    t = 0.0
    for r in range(3):
        for (f, n) in groups:
            t += float(rng.exponential(2e-5))
            reqs.append((t, f, n))
    for streams, serial in [(4, False), (1, False), (1, True)]:
        outs = [np.zeros_like(a) for a in alone]
        _say(f"serve {len(reqs)} requests streams={streams} serial={serial}")
        stats, lat = gpu.serve(imgs, reqs, streams=streams, serial_plain=serial, keep_outputs=outs)
        assert stats["completed"] == len(reqs) and stats["rejected"] == 0
        assert (lat > 0).all()
        for i, (a, b) in enumerate(zip(alone, outs)):
            assert np.array_equal(a, b), (streams, serial, i)
        if not serial and streams == 4:
            assert stats["batches"] < len(reqs), "requests that arrive together are batched"
    imgs.free()


def test_serve_budget_queues_and_rejects(gpu):
    items = small_batch_items(22, 12, max_k=7, serve_mix=False)
    imgs = gpu.DeviceImages(items)
    groups = _texture_requests(items)
    need = []
    for (f, n) in groups:
        b = 0
        for it in items[f:f + n]:
            nb = ((it["width"] + 3) // 4) * ((it["height"] + 3) // 4) * (8 if it["dxt"] == 0 else 16)
            b += (nb + 255) // 256 * 256 + it["width"] * it["height"] * 4
        need.append(b)
    budget = max(need) + 1
    reqs = [(0.0, f, n) for (f, n) in groups]
    _say(f"serve with a budget of one request ({budget} bytes)")
    stats, lat = gpu.serve(imgs, reqs, hbm_budget=budget, streams=4)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert stats["max_inflight"] < len(reqs)
    small = min(need)
    _say(f"serve with a budget below the largest request ({small} bytes)")
    stats, lat = gpu.serve(imgs, reqs, hbm_budget=small, streams=2)
    assert stats["rejected"] >= 1
    assert stats["completed"] + stats["rejected"] == len(reqs)
    assert ((lat < 0) == np.array([n > small for n in need])).all()
    imgs.free()


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def serve_pool(gpu):
    """The serving pool on the device: 256 textures of the serving mix (70%
    quick mode, 30% nrandom 64 + LOOP), base levels uploaded and every mip
    level built on the device, as the benchmark's serving leg does."""
    from s2tc_amd import workloads
    exp = json.load(open(os.path.join(ROOT, "tests", "golden", "digests.json")))["workloads"]["serve_pool"]
    _say("building the serving pool (256 textures, device mips)")
    textures = workloads.batch_textures(workloads.SERVE_SEED, workloads.SERVE_POOL, serve_mix=True)
    imgs = gpu.DeviceImages.from_textures(textures)
    groups, f = [], 0
    for t in textures:
        n = max(t["width"], t["height"]).bit_length()   # levels down to 1x1
        groups.append((f, n))
        f += n
    assert f == len(imgs.items) == exp["jobs"] and imgs.texels == exp["texels"]
    yield imgs, groups, exp["texture_sha256"]
    imgs.free()


def _poisson_requests(gpu, imgs, groups, rate, duration, seed):
    """Requests drawn from the pool with exponential inter-arrival times; each
    request gets its own copies of its texture's jobs (same device sources),
    so every request's blocks come back to a buffer of their own."""
    rng = np.random.default_rng(seed)
    n = max(1, int(rate * duration))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    picks = rng.integers(0, len(groups), n)
    total = sum(groups[p][1] for p in picks)
    jobs = (gpu.Job * total)()
    reqs, outs, k = [], [], 0
    for a, p in zip(arrivals, picks):
        f, m = groups[p]
        for q in range(m):
            j = imgs.jobs[f + q]
            jobs[k + q] = j
            outs.append(np.zeros(gpu.nblocks(j.width, j.height) * gpu.block_bytes(j.dxt), np.uint8))
        reqs.append((float(a), k, m))
        k += m
    return types.SimpleNamespace(jobs=jobs), reqs, outs, picks


@pytest.mark.parametrize("rate,duration,serial", [(10000, 0.2, False), (50000, 0.05, False), (2000, 0.1, True)])
def test_serve_poisson_matches_reference_digests(gpu, serve_pool, rate, duration, serial):
    """A Poisson stream from the serving pool through s2tc_amd_serve (10k
    req/s, the benchmark's middle rate; 50k req/s, where requests queue and
    batches are large; serial plain launches at 2k req/s): every request's
    blocks, all mip levels, hash to the compiled reference's digest of its
    texture."""
    imgs, groups, digests = serve_pool
    view, reqs, outs, picks = _poisson_requests(gpu, imgs, groups, rate, duration, seed=rate + int(serial))
    _say(f"serve {len(reqs)} Poisson requests at {rate}/s serial={serial} ({len(outs)} images)")
    stats, lat = gpu.serve(view, reqs, hbm_budget=64 << 30, streams=4, max_batch=256, serial_plain=serial,
                           keep_outputs=outs)
    _say(f"served: p50 {stats['p50_ms']:.3f} ms p99 {stats['p99_ms']:.3f} ms batches {stats['batches']}")
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert (lat > 0).all()
    bad = []
    for r, ((_, first, m), p) in enumerate(zip(reqs, picks)):
        h = hashlib.sha256(b"".join(outs[first + q].tobytes() for q in range(m))).hexdigest()
        if h != digests[p]:
            bad.append((r, int(p)))
    assert not bad, f"{len(bad)} of {len(reqs)} requests differ from the reference (request, texture): {bad[:10]}"
    if not serial:
        assert stats["batches"] < len(reqs), "requests that arrive while others run are batched"
