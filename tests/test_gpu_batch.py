"""Batched work pool: every scene's film equals rendering it alone, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven(rp, gpu_available):
    items = []
    specs = [(rp.SYNTH_BATCH, 40, 32, 3, dict(target_triangles=6000)),
             (rp.SYNTH_BATCH, 64, 48, 1, dict(target_triangles=60000)),
             (rp.SYNTH_CORNELL, 32, 32, 5, {}),
             (rp.SYNTH_INVERSE, 24, 24, 2, {}),
             (rp.SYNTH_BATCH, 16, 16, 7, dict(target_triangles=20000))]
    keep = []
    for i, (kind, w, h, spp, opts) in enumerate(specs):
        s = rp.Synthetic(kind, seed=100 + i, width=w, height=h, **opts)
        scene = rp.Scene(s.desc)
        keep.append(s)
        items.append((scene, spp, 1 + i))
    return keep, items


def _alone(rp, items):
    return [scene.render_frame(rp.frame_config(spp, mode=rp.MODE_MEGAKERNEL, seed=seed))[0]
            for scene, spp, seed in items]


@pytest.mark.parametrize("cfg", [
    dict(scheduler=0, group_size=128),
    dict(scheduler=0, group_size=64),
    dict(scheduler=1, capacity=1 << 12, group_size=128, sort_bits=0, use_graph=0),
    dict(scheduler=1, capacity=1 << 12, group_size=256, sort_bits=2, use_graph=1),
    dict(scheduler=1, capacity=1 << 16, group_size=64, sort_bits=4, use_graph=1),
    dict(scheduler=1, capacity=1 << 12, group_size=128, use_graph=1, flags=1),
    dict(scheduler=1, capacity=1 << 12, group_size=128, use_graph=1, tail_paths=256),
], ids=["persistent128", "persistent64", "pool-nograph", "pool-graph-sort2", "pool-graph-sort4", "pool-no-tail",
        "pool-tail-256"])
def test_batch_films_equal_single_frames(rp, uneven, cfg):
    _, items = uneven
    batch = rp.Batch(items)
    films, st = batch.render(rp.Batch.config(**cfg))
    print(f"[batch] {cfg}: {st}", flush=True)
    assert st["failed_paths"] == 0
    assert st["paths"] == sum(s.width * s.height * spp for s, spp, _ in items)
    for got, want in zip(films, _alone(rp, items)):
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_batch_graph_replays_repeatable(rp, uneven):
    _, items = uneven
    batch = rp.Batch(items)
    cfg = rp.Batch.config(scheduler=1, capacity=1 << 10, group_size=128, use_graph=1)
    a, sa = batch.render(cfg)
    b, sb = batch.render(cfg)
    assert sa["graph_launches"] >= 1
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_autotune_returns_valid_config(rp, uneven):
    _, items = uneven
    batch = rp.Batch(items)
    cfg, sec = batch.autotune()
    print(f"[autotune] {rp.config_dict(cfg)} {sec * 1e3:.2f} ms", flush=True)
    assert sec > 0 and cfg.scheduler in (0, 1) and cfg.group_size in (64, 128, 256)
    films, _ = batch.render(cfg)
    for got, want in zip(films, _alone(rp, items)):
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_pool_tail_keeps_ray_counts(rp, uneven):
    """The pool's latency tail (every path issued, few live) traces the same
    rays as bounce iterations to the end and as the scenes alone."""
    _, items = uneven
    batch = rp.Batch(items)
    alone = [scene.render_frame(rp.frame_config(spp, mode=rp.MODE_MEGAKERNEL, seed=seed))[1]
             for scene, spp, seed in items]
    want = (sum(a["extension_rays"] for a in alone), sum(a["shadow_rays"] for a in alone))
    for kw in (dict(flags=rp.BATCH_NO_TAIL), dict(), dict(tail_paths=256)):
        _, st = batch.render(rp.Batch.config(scheduler=1, capacity=1 << 12, group_size=128, **kw))
        print(f"[pool tail] {kw}: {st}", flush=True)
        assert (st["extension_rays"], st["shadow_rays"]) == want


@pytest.fixture(scope="module")
def mix64(rp, gpu_available):
    """The 64-scene uneven mix (5k-500k triangles, spp 4/16/64, depth 8) at
    one eighth of its resolutions (16-128 px squares)."""
    mix = rp.batch_mix(64, resolution_divisor=8)
    items = [(rp.Scene(s.desc), spp, seed) for s, spp, seed in mix]
    return mix, items


def test_mix64_pool_films_equal_scenes_alone(rp, mix64):
    mix, items = mix64
    batch = rp.Batch(items)
    cfg, _ = batch.autotune()
    print(f"[mix64] autotuned {rp.config_dict(cfg)}", flush=True)
    for c in (cfg, rp.Batch.config(scheduler=rp.SCHED_PERSISTENT, group_size=128)):
        films, st = batch.render(c)
        print(f"[mix64] {rp.config_dict(c)}: {st}", flush=True)
        assert st["failed_paths"] == 0
        assert st["paths"] == sum(s.width * s.height * spp for s, spp, _ in items)
        for i, (got, want) in enumerate(zip(films, _alone(rp, items))):
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"scene {i}"


def test_mix64_records_match_oracle(rp, orc, mix64):
    """Records of a subset of the mix (the smallest, the largest and two
    middle scenes by triangle count) against the oracle."""
    from tests.parity_util import AGREE_FRACTION, agreement

    mix, items = mix64
    order = np.argsort([s.d.triangle_count for s, _, _ in mix])
    for i in (order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]):
        synth, spp, seed = mix[i]
        scene = items[i][0]
        n = scene.width * scene.height
        req = rp.request(seed=seed, pixel_begin=0, pixel_count=min(n, 2048), sample_count=2)
        print(f"[mix64] scene {i}: {synth.d.triangle_count} triangles", flush=True)
        gpu = scene.sample(req)
        cpu = orc.scene(synth.desc).sample(req)
        assert np.array_equal(gpu["sobol_index"], cpu["sobol_index"])
        assert np.array_equal(gpu["primary_triangle"], cpu["primary_triangle"])
        assert agreement(gpu["xyz"], cpu["xyz"]) >= AGREE_FRACTION
