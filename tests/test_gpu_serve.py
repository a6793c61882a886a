"""The serving loop: every request's blocks are the same alone and in the
mix; budget admission queues or rejects instead of failing."""
import sys

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
